"""Device engine behind TimeEvolutionProblem.compute(): one libjja plan plus the
device-resident state, inputs and outputs of one time evolution.

``Engine(problem, th_mask, I_mask)`` performs the reference's set-up
(time_evolution.py:466-521) on the host and uploads everything once;
``advance(n)`` integrates n further steps with persistent kernel launches;
``results()`` downloads the stored steps in the reference's layout and the vortex
counts computed on the device. Device memory is owned by torch tensors.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _native
from .circuit import device_cpr_terms
from .solver_plan import CircuitPlan


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pyjjasim_amd: the time evolution needs a HIP (ROCm) device; "
                           "none is visible")
    return torch


def source_layout(value, n, W, Nt):
    """Device layout of an (n, W, Nt)-broadcastable input: the un-broadcast
    elements in [t][w][i] order (element index fastest) and the element strides
    (s_i, s_w, s_t), 0 on broadcast axes: element (i, w, t) is
    flat[i*s_i + w*s_w + t*s_t]."""
    view = np.broadcast_to(np.asarray(value, dtype=np.float64), (n, W, Nt))
    keep = [st != 0 and d > 1 for d, st in zip(view.shape, view.strides)]
    sub = view[tuple(slice(None) if k else slice(0, 1) for k in keep)]
    dev = np.ascontiguousarray(sub.transpose(2, 1, 0))
    ni, nw = dev.shape[2], dev.shape[1]
    strides = (1 if keep[0] else 0, ni if keep[1] else 0, ni * nw if keep[2] else 0)
    return dev.reshape(dev.shape[0], nw, ni).reshape(-1).copy(), strides


class DeviceSource:
    """Device copy of one input (f, Is, Vs or T) plus its jja_source descriptor.

    Arrays are uploaded un-broadcast and reordered to [t][w][i] (element index
    fastest), so that consecutive threads read consecutive elements; a broadcast
    axis keeps stride 0 and takes no memory. Callables are evaluated on the host
    for each launch's chunk of steps.
    """

    def __init__(self, torch, value, n, W, Nt, zero):
        self.callable = callable(value)
        self.value = value
        self.n, self.W = n, W
        self.src = _native.Source()
        self.src.zero = int(bool(zero))
        self.tensor = None
        if not self.callable:
            sep = hasattr(value, "amplitude") and hasattr(value, "base")
            base = value.base[:, :, None] if sep else value
            dev, (s_i, s_w, s_t) = source_layout(base, n, W, Nt)
            if dev.size == 0:                  # no elements (no faces): a valid base, never read
                dev = np.zeros(1)
            self.tensor = torch.from_numpy(dev).cuda()
            self.src.base = self.tensor.data_ptr()
            self.src.s_i, self.src.s_w, self.src.s_t = s_i, s_w, s_t
            self.src.t0 = 0
            # read every step, per column: a term that varies over the columns
            # (a term shared by all columns stays cache-resident, like the
            # junction coefficients, and costs no HBM traffic per column)
            self.bytes_per_step = 8 * (n if s_i else 1) * W if s_w else 0
            if sep:
                # amplitude[w, t] as [t][w]: value = base * amplitude, one product
                amp, (a_1, a_w, a_t) = source_layout(value.amplitude[None, :, :], 1, W, Nt)
                self.amp_tensor = torch.from_numpy(amp).cuda()
                self.src.amp = self.amp_tensor.data_ptr()
                self.src.a_w, self.src.a_t = a_w, a_t
        else:
            self.bytes_per_step = 8 * n * W

    def fill_chunk(self, torch, t0, n_steps):
        if not self.callable:
            return
        buf = np.empty((n_steps, self.W, self.n))
        for k in range(n_steps):
            buf[k] = np.broadcast_to(self.value(t0 + k), (self.n, self.W)).T
        flat = buf.reshape(-1) if buf.size else np.zeros(1)   # no elements: a valid base
        self.tensor = torch.from_numpy(flat).cuda()
        self.src.base = self.tensor.data_ptr()
        self.src.s_i, self.src.s_w, self.src.s_t = 1, self.n, self.n * self.W
        self.src.t0 = t0


def reference_noise(Nj, W):
    """The reference's legacy-RNG draw order (time_evolution.py:533-537)."""
    state = {}

    def draw(i):
        if Nj > 500:
            if i % 3 == 0:
                state["r"] = np.random.randn(Nj, W)
            else:
                state["r"] = state["r"][np.random.permutation(Nj), :]
        else:
            state["r"] = np.random.randn(Nj, W)
        return state["r"]
    return draw


# stored steps beyond this many bytes stream to the host during the run
DEFAULT_RESIDENT_OUTPUT_BYTES = 16 << 30


def set_cpr(desc, cpr):
    """The current-phase relation's sine series in a problem descriptor."""
    terms, coef = cpr
    desc.cpr_terms = int(terms)
    for k in range(len(coef)):
        desc.cpr_coef[k] = float(coef[k])


def default_block_size(Nj, Nf):
    work = max(Nj, Nf)
    for bs in (64, 128, 256, 512):
        if work <= 2 * bs:
            return bs
    return 1024


def device_compute_units():
    return _torch().cuda.get_device_properties(0).multi_processor_count


def auto_parts(host, columns, stencil_width, max_parts=16, junctions_per_part=24576):
    """Workgroups per column for a plan: 1 for an LDS face store, else the
    largest power of two up to max_parts whose workgroups, one per compute
    unit, all fit on the device and each keep at least junctions_per_part
    junctions (below that the column barriers cost more than the split saves;
    SquareArray(512,512) at W=16 takes 16 parts: 1.75 ms/step against 2.03 at 8)."""
    if not host.global_store or stencil_width != 3 or host.block_size not in (256, 512, 1024):
        return 1
    n_cu = device_compute_units()
    parts = 1
    while (parts < max_parts and 2 * parts * columns <= n_cu
           and 2 * parts * junctions_per_part <= host.Nj):
        parts *= 2
    return parts


class Engine:
    """Device state and launches of one problem. With ``shared = (handle, index,
    host)`` the problem is problem ``index`` of a plan that a BatchEngine owns."""

    def __init__(self, problem, th_mask, I_mask, shared=None):
        cpr = device_cpr_terms(problem.current_phase_relation)
        if cpr is None:
            raise ValueError("the GPU time evolution evaluates the default current-phase relation "
                             "Ic*sin(theta) and sine series (FourierCPR); got another CurrentPhaseRelation")
        torch = self.torch = _torch()
        self.problem = problem
        circuit = problem.circuit
        Nj, Nf, W, Nt = circuit._Nj(), circuit._Nf(), problem.problem_count, problem._Nt()
        self.Nj, self.Nf, self.W, self.Nt = Nj, Nf, W, Nt
        dt = problem._dt()
        sw = problem.stencil_width
        S = self.S = sw - 1
        self.lib = lib = _native.lib()
        if shared is None:
            self.block_size = problem.block_size or default_block_size(Nj, Nf)
            self.host = CircuitPlan(circuit, dt, sw, leaf_size=problem.leaf_size,
                                    block_size=self.block_size,
                                    face_store=getattr(problem, "face_store", "auto"),
                                    local_stages=getattr(problem, "local_stages", "auto"),
                                    solver=getattr(problem, "solver", "auto"))
            parts = getattr(problem, "parts", None)
            if parts is None:
                parts = auto_parts(self.host, W, sw)
            if parts > 1:
                self.host = self.host.relayout(self.block_size, parts)
            if getattr(problem, "column_pairs", False):
                if self.host.uses_lattice and getattr(problem, "solver", "auto") == "auto":
                    self.host = self.host.relayout(self.block_size, solver="cholesky")
                self.host = self.host.paired()
                self.block_size = self.host.block_size
            self._desc = self.host.desc(max_columns=W)
            set_cpr(self._desc, cpr)
            self.handle = C.c_void_p()
            _native.check(lib.jja_plan_create(1, C.byref(self._desc), C.byref(self.handle)))
            self.owns_plan, self.index = True, 0
        else:
            self.handle, self.index, self.host = shared
            self.block_size = self.host.block_size
            self.owns_plan = False
        zd = lib.jja_plan_zbuf_doubles(self.handle, self.index)
        if zd == 0 and lib.jja_plan_lds_bytes(self.handle, self.index) == 0 and Nf:
            self.close()
            raise _native.NativeError(
                f"circuit with {Nf} faces needs {self.host.lds_bytes()} bytes of LDS per "
                "column; larger than this device's workgroup LDS")
        dev = self.dev = torch.device("cuda")
        # global face store: one array of Nf + stage-scratch doubles per column
        self.zbuf = torch.empty((W, zd), dtype=torch.float64, device=dev) if zd else None
        f64 = torch.float64
        cfgs = [problem.config_at_minus_1, problem.config_at_minus_2]
        if sw >= 4:
            cfgs.append(problem.config_at_minus_3)
        if sw >= 5:
            cfgs.append(problem.config_at_minus_4)
        self.cfgs = cfgs
        hist = np.zeros((S, W, Nj))
        for k, cfg in enumerate(cfgs):          # theta at step m sits in slot m mod S
            hist[(-(k + 1)) % S] = np.asarray(cfg, dtype=np.float64).reshape(Nj, W).T
        self.theta = torch.from_numpy(hist).to(dev)
        self.srcs = {}
        for name, val, n, td in (("Is", problem.current_sources, Nj, problem._Is_is_timedep),
                                 ("Vs", problem.voltage_sources, Nj, problem._Vs_is_timedep),
                                 ("f", problem.external_flux, Nf, problem._f_is_timedep),
                                 ("T", problem.temperature, Nj, problem._T_is_timedep)):
            sample = np.broadcast_to(val(0), (n, W)) if callable(val) else val[:, :, 0]
            zero = (not td) and bool(np.allclose(sample, 0))       # time_evolution.py:509-519
            self.srcs[name] = DeviceSource(torch, problem._raw[name], n, W, Nt, zero)
        self.theta_s = torch.zeros((W, Nj), dtype=f64, device=dev) \
            if not self.srcs["Vs"].src.zero else None
        # noise amplitude sqrt(2 T Rv), formed once when T is constant in time
        # (the same expression as time_evolution.py:538, correctly rounded sqrt)
        self.sigma = None
        T_src = self.srcs["T"]
        if not T_src.src.zero and not T_src.callable and not problem._T_is_timedep:
            Traw = np.broadcast_to(np.array(problem._raw["T"], dtype=np.float64), (Nj, W, Nt))[:, :, 0]
            sig = np.sqrt(2.0 * Traw * self.host.Rv[:, None])
            if not np.ptp(sig):
                sig = sig[:1, :1]                 # one value: every junction reads one address
            elif not np.ptp(sig, axis=1).any():
                sig = sig[:, :1]
            self.sigma = DeviceSource(torch, sig[:, :, None], Nj, W, Nt, False)
        # u = x/c0 - theta_s between the junction pass and the face phase: an
        # LDS face store's scratch (a global store's face phase forms it from x)
        self.u = None if self.host.global_store else torch.zeros((W, Nj), dtype=f64, device=dev)
        self.th_mask, self.I_mask = th_mask, I_mask
        self.n_th, self.n_I = int(th_mask.sum()), int(I_mask.sum())
        # stored steps: all resident on the device, or (when they exceed the
        # budget) streamed to the host through rings of output slots, each
        # launch's slots copied out after it
        budget = getattr(problem, "max_resident_output_bytes", None)
        budget = DEFAULT_RESIDENT_OUTPUT_BYTES if budget is None else int(budget)
        row = 8 * W * Nj
        self.streaming = shared is None and (self.n_th + self.n_I) * row > budget
        if self.streaming:
            self.ring = max(1, budget // (2 * row))
            self.th_host = np.empty((self.n_th, W, Nj))
            self.I_host = np.empty((self.n_I, W, Nj))
            self.n_host = np.empty((self.n_th, W, Nf), dtype=np.int32) if Nf else None
            cap_th, cap_I = min(self.n_th, self.ring), min(self.n_I, self.ring)
        else:
            self.ring = max(self.n_th, self.n_I, 1)
            cap_th, cap_I = self.n_th, self.n_I
        self._th_rank = np.cumsum(th_mask)        # stored theta steps up to and including t
        self._I_rank = np.cumsum(I_mask)
        th_slot = np.where(th_mask, (self._th_rank - 1) % self.ring, -1).astype(np.int32)
        I_slot = np.where(I_mask, (self._I_rank - 1) % self.ring, -1).astype(np.int32)
        self.th_slot = torch.from_numpy(th_slot).to(dev)
        self.I_slot = torch.from_numpy(I_slot).to(dev)
        self.th_out = torch.empty((max(cap_th, 1), W, Nj), dtype=f64, device=dev)
        self.I_out = torch.empty((max(cap_I, 1), W, Nj), dtype=f64, device=dev)
        run = self.run = _native.ProblemRun()
        run.W = W
        run.column_offset = 0
        run.dt = dt
        run.theta = self.theta.data_ptr()
        run.theta_s = self.theta_s.data_ptr() if self.theta_s is not None else None
        run.u = self.u.data_ptr() if self.u is not None else None
        if self.sigma is not None:
            run.sigma = self.sigma.src
        run.th_slot = self.th_slot.data_ptr()
        run.I_slot = self.I_slot.data_ptr()
        run.slot_t0 = 0
        run.th_out = self.th_out.data_ptr()
        run.I_out = self.I_out.data_ptr()
        run.zbuf = self.zbuf.data_ptr() if self.zbuf is not None else None
        T_zero = self.srcs["T"].src.zero
        self.ref_noise = (not T_zero) and problem.noise == "reference"
        if not T_zero and problem.noise == "philox":
            run.seed = int(problem.seed) if problem.seed is not None \
                else int(np.random.randint(0, 2 ** 62, dtype=np.int64))
        self.draw = reference_noise(Nj, W) if self.ref_noise else None
        self.host_chunks = self.ref_noise or any(s.callable for s in self.srcs.values())
        spl = problem.steps_per_launch
        self.steps_per_launch = spl if spl else (500 if self.host_chunks else 2000)
        self.t = 0
        self._noise_buf = None

    def prepare(self, n_steps):
        """Host work before a launch of n_steps steps: callable inputs and the
        reference-noise draws for the chunk; returns the run descriptor."""
        torch = self.torch
        for s in self.srcs.values():
            s.fill_chunk(torch, self.t, n_steps)
        if self.ref_noise:
            z = np.empty((n_steps, self.W, self.Nj))
            for k in range(n_steps):
                z[k] = self.draw(self.t + k).T
            self._noise_buf = torch.from_numpy(z.reshape(-1)).to(self.dev)
            self.run.noise = self._noise_buf.data_ptr()
            self.run.noise_t0 = self.t
        run = self.run
        run.Is, run.Vs, run.f, run.T = (self.srcs[k].src for k in ("Is", "Vs", "f", "T"))
        return run

    def launch(self, n_steps, stream=None):
        """One persistent launch of n_steps steps starting at the current step
        (within time_step_count: the sources and output-slot maps end there)."""
        if n_steps < 0 or self.t + n_steps > self.Nt:
            raise ValueError("launch beyond time_step_count")
        torch = self.torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        run = self.prepare(n_steps)
        _native.check(self.lib.jja_evolve(self.handle, C.byref(run), n_steps, self.t, stream))
        if self.host_chunks:
            torch.cuda.current_stream().synchronize()      # chunk buffers are reused
        self.t += n_steps

    def _stored_between(self, rank, t0, t1):
        """Ranks [r0, r1) of the stored steps in [t0, t1)."""
        return (int(rank[t0 - 1]) if t0 > 0 else 0), (int(rank[t1 - 1]) if t1 > 0 else 0)

    def _ring_limit(self, end):
        """The longest launch from the current step (at most to `end`) whose
        stored steps each fit in their ring without overwriting a slot that has
        not been copied out."""
        n = end - self.t
        for rank in (self._th_rank, self._I_rank):
            r0 = int(rank[self.t - 1]) if self.t > 0 else 0
            # last step whose rank stays within r0 + ring
            k = int(np.searchsorted(rank[self.t:end], r0 + self.ring, side="right"))
            n = min(n, max(k, 1))
        return n

    def _flush(self, t0, t1):
        """Copy the stored steps of [t0, t1) from their ring slots to the host
        arrays, with their vortex counts (computed on the device first)."""
        torch = self.torch
        stream = torch.cuda.current_stream().cuda_stream
        W = self.W
        for rank, dev_out, host, with_n in ((self._th_rank, self.th_out, self.th_host, True),
                                            (self._I_rank, self.I_out, self.I_host, False)):
            r0, r1 = self._stored_between(rank, t0, t1)
            r = r0
            while r < r1:                       # ring segments [s0, s0 + cnt) without wrap-around
                s0 = r % self.ring
                cnt = min(r1 - r, self.ring - s0)
                if with_n and self.n_host is not None:
                    n_dev = torch.empty((cnt, W, self.Nf), dtype=torch.int32, device=self.dev)
                    _native.check(self.lib.jja_vortex(self.handle, self.index, dev_out[s0].data_ptr(),
                                                      cnt * W, n_dev.data_ptr(), stream))
                    self.n_host[r:r + cnt] = n_dev.cpu().numpy()
                host[r:r + cnt] = dev_out[s0:s0 + cnt].cpu().numpy()
                r += cnt

    def advance(self, n_steps):
        if self.t + n_steps > self.Nt:
            raise ValueError("advance beyond time_step_count")
        end = self.t + n_steps
        while self.t < end:
            n = min(self.steps_per_launch, end - self.t)
            if self.streaming:
                n = self._ring_limit(self.t + n)
            t0 = self.t
            self.launch(n)
            if self.streaming:
                self._flush(t0, self.t)

    def results(self):
        torch = self.torch
        Nj, W, S = self.Nj, self.W, self.S
        stream = torch.cuda.current_stream().cuda_stream
        n_dev = None
        if self.n_th and self.Nf and not self.streaming:
            n_dev = torch.empty((self.n_th, W, self.Nf), dtype=torch.int32, device=self.dev)
            _native.check(self.lib.jja_vortex(self.handle, self.index, self.th_out.data_ptr(),
                                              self.n_th * W, n_dev.data_ptr(), stream))
        torch.cuda.synchronize()
        _native.check(self.lib.jja_plan_status(self.handle))
        if self.streaming:
            th_s, I_s = self.th_host, self.I_host
        else:
            th_s = self.th_out[:self.n_th].cpu().numpy() if self.n_th else None
            I_s = self.I_out[:self.n_I].cpu().numpy() if self.n_I else None
        th = th_s.transpose(2, 1, 0) if self.n_th else np.zeros((Nj, W, 0))
        I = I_s.transpose(2, 1, 0) if self.n_I else np.zeros((Nj, W, 0))
        n = None
        if self.streaming and self.n_th and self.Nf:
            n = self.n_host.astype(np.int64).transpose(2, 1, 0)
        elif n_dev is not None:
            n = n_dev.cpu().numpy().astype(np.int64).transpose(2, 1, 0)
        elif self.n_th:
            n = np.zeros((0, W, self.n_th), dtype=np.int64)
        # prefix of initial conditions (time_evolution.py:484-502)
        pre = np.stack([np.asarray(self.cfgs[S - 1 - k]).reshape(Nj, W) for k in range(S)], axis=2)
        Ic = self.problem.circuit._Ic()[:, None]
        cp = self.problem.current_phase_relation
        pre_I = np.stack([cp.eval(Ic, pre[:, :, k]) for k in range(S)], axis=2)
        return (np.concatenate((pre, th), axis=2), np.concatenate((pre_I, I), axis=2), n)

    def bytes_per_step(self):
        """Algorithmic HBM bytes of one time step of the whole batch, without
        stored outputs (bytes_per_stored_step adds those):
        W (24 Nj + 48 Nf) + 24 nnz + 8 Nj k_par + sources.
        24 Nj: theta_{t-1}, theta_{t-2} read and theta_t written; 48 Nf: face
        right-hand side written, read and written by each triangular pass and
        read by the A^T gather; nnz: factor entries one triangular pass streams
        (off-block rows of L plus the dense diagonal-block inverses), each pass
        reading an 8-byte value and a 4-byte index; k_par = 5 coefficients per
        junction (Ic, c0, c1, c2, noise scale); sources: terms that vary over
        the columns, 8 n W bytes each. With the lattice solve the factor term is
        the solve's tables instead (half the sine matrix and the pivots)."""
        if self.host.uses_lattice:
            solve = self.host.lattice.table_bytes()
        else:
            sp = self.host.solve_plan
            solve = 24 * (len(sp.lo_val) + len(sp.inv_cm))
        src = sum(s.bytes_per_step for k, s in self.srcs.items() if not s.src.zero)
        return self.W * (24 * self.Nj + 48 * self.Nf) + solve + 8 * self.Nj * 5 + src

    def bytes_per_stored_step(self):
        """Output bytes of a step whose theta and current are both stored
        (8 Nj W each); a step with one of them stores half."""
        return 8 * self.Nj * self.W

    def close(self):
        if getattr(self, "owns_plan", False) and getattr(self, "handle", None) is not None \
                and self.handle.value:
            self.lib.jja_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        # at interpreter exit the HIP runtime's own exit handlers may already
        # have run: leave the memory to the process teardown instead of calling
        # into the runtime (close() explicitly to free it earlier)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class BatchEngine:
    """Several problems (any circuits and column counts, one stencil width and
    step count) in one plan: every launch advances all their columns together,
    each column worked by its problem's part count of workgroups
    (time_evolution.py:461-582 for each problem). ``hosts`` are prepared plans
    (CircuitPlan, possibly with parts > 1); otherwise plans are built here."""

    def __init__(self, problems, masks, block_size=None, hosts=None):
        if not problems:
            raise ValueError("empty batch")
        Nt = problems[0]._Nt()
        sw = problems[0].stencil_width
        if any(p._Nt() != Nt or p.stencil_width != sw for p in problems):
            raise ValueError("a batch needs one time_step_count and one stencil width")
        self.torch = _torch()
        self.lib = _native.lib()
        if hosts is not None:
            self.block_size = hosts[0].block_size
            self.hosts = list(hosts)
        else:
            self.block_size = block_size or max(default_block_size(p.circuit._Nj(), p.circuit._Nf())
                                                for p in problems)
            self.hosts = [CircuitPlan(p.circuit, p._dt(), sw, leaf_size=p.leaf_size,
                                      block_size=self.block_size,
                                      face_store=getattr(p, "face_store", "auto"),
                                      local_stages=getattr(p, "local_stages", "auto"),
                                      solver=getattr(p, "solver", "auto")) for p in problems]
        descs = (_native.ProblemDesc * len(problems))(*[h.desc(max_columns=p.problem_count)
                                                         for h, p in zip(self.hosts, problems)])
        for d, p in zip(descs, problems):
            cpr = device_cpr_terms(p.current_phase_relation)
            if cpr is None:
                raise ValueError("the GPU time evolution evaluates the default current-phase relation "
                                 "Ic*sin(theta) and sine series (FourierCPR); got another CurrentPhaseRelation")
            set_cpr(d, cpr)
        self._descs = descs
        self.handle = C.c_void_p()
        _native.check(self.lib.jja_plan_create(len(problems), descs, C.byref(self.handle)))
        self.engines = []
        try:
            for i, (p, (thm, Im)) in enumerate(zip(problems, masks)):
                e = Engine(p, thm, Im, shared=(self.handle, i, self.hosts[i]))
                self.engines.append(e)
        except Exception:
            self.close()
            raise
        self.Nt = Nt
        self.t = 0
        spl = [p.steps_per_launch for p in problems if p.steps_per_launch]
        self.steps_per_launch = min(spl) if spl else min(e.steps_per_launch for e in self.engines)
        self.host_chunks = any(e.host_chunks for e in self.engines)

    def launch(self, n_steps, stream=None):
        if n_steps < 0 or self.t + n_steps > self.Nt:
            raise ValueError("launch beyond time_step_count")
        torch = self.torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        runs = (_native.ProblemRun * len(self.engines))(*[e.prepare(n_steps) for e in self.engines])
        _native.check(self.lib.jja_evolve(self.handle, runs, n_steps, self.t, stream))
        if self.host_chunks:
            torch.cuda.current_stream().synchronize()
        self.t += n_steps
        for e in self.engines:
            e.t = self.t

    def advance(self, n_steps):
        if self.t + n_steps > self.Nt:
            raise ValueError("advance beyond time_step_count")
        end = self.t + n_steps
        while self.t < end:
            self.launch(min(self.steps_per_launch, end - self.t))

    def results(self):
        return [e.results() for e in self.engines]

    def bytes_per_step(self):
        return sum(e.bytes_per_step() for e in self.engines)

    def close(self):
        for e in self.engines:
            e.close()
        if self.handle is not None and self.handle.value:
            self.lib.jja_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        # at interpreter exit the HIP runtime's own exit handlers may already
        # have run: leave the memory to the process teardown instead of calling
        # into the runtime (close() explicitly to free it earlier)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
