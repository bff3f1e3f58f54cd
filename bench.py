"""Benchmark: batched IV sweep on one MI355X, FP64.

Workload: SquareArray(100,100) (19,800 junctions, 9,801 faces), per-junction Ic
log-normal (sigma 0.1, seeded), beta_C = 1, f = 0.25, W = 256 bias currents
linspace(0, 2.0) * current_base(0), T = 0.01 with the seeded in-kernel Gaussian
stream, dt = 0.05, theta stored at steps Nt//3 and Nt-1.

A bench "step" is one time step of all 256 columns. The timed region covers
``--steps`` steps (persistent kernel launches) with all inputs resident in HBM,
after ``--warmup`` untimed steps. The plain loop runs between the batched
run's host set-up and its warmup (its GPU work also brings the device to its
working clock). Reported:

  value      junction-steps/s = Nj * W * steps / wall time
  roofline   algorithmic HBM bytes per launch / mean launch time (HIP events on the
             launch stream) against the 8 TB/s HBM peak
  plain      the one-problem-at-a-time GPU loop (W = 1) on the same circuit
  cpu_baseline  the FP64 numpy/scipy restatement of the reference scheme (same
             library calls), timed on this host on a bounded sample of steps

Usage: python bench.py [--gpus 1] [--steps 20000] [--warmup 1000] [--plain-steps 0]
       [--no-cpu] [--no-adjoint] [--copy-calibration MIB]
       python bench.py --workload uneven_batch [--steps 10000] [--warmup 200] [--problems 64]
         (64 seeded problems of uneven size, one launch per chunk, autotuned; its own
          JSON line)
       python bench.py --workload large --steps 200 [--checkpoint 1000]
         (SquareArray(512,512) W=16 forward + adjoint d loss/d f on a bounded sample of
          steps; its own JSON line)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "batched_iv_sweep_square100_W256"


def bias_amplitudes(W):
    """Bias points of the IV sweep, in units of current_base(0)."""
    return np.linspace(0.0, 2.0, W)


def make_problem(pa, n, W, Nt, T=0.01, seed=12345, steps_per_launch=None, block_size=None,
                 amplitudes=None, **engine_options):
    """The IV sweep: the bias is current_base(0) times one amplitude per column,
    passed as a SeparableSource (the device forms the product, bit for bit the
    reference's current_base[:, None, None] * linspace[None, :, None])."""
    c = pa.SquareArray(n, n)
    rng = np.random.default_rng(20261202)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    c.set_capacitance(1.0)
    amp = bias_amplitudes(W) if amplitudes is None else np.asarray(amplitudes, dtype=np.float64)
    Is = pa.SeparableSource(c.current_base(0.0), amp[:, None])
    prob = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=Nt, external_flux=0.25,
                                   current_sources=Is, temperature=T,
                                   store_time_steps=[Nt // 3, Nt - 1], store_current=False,
                                   store_voltage=False, seed=seed,
                                   steps_per_launch=steps_per_launch, block_size=block_size,
                                   **engine_options)
    return c, prob


def run_gpu(pa, torch, n, W, steps, warmup, steps_per_launch, block_size, events=True, amplitudes=None,
            prelude=None, download=True, **engine_options):
    """The timed batched run. ``prelude`` (GPU work of its own, such as the
    plain loop) runs after the host has built this run's plan and before the
    warmup steps, so that no host set-up lies between it and the timed steps."""
    from pyjjasim_amd.engine import Engine
    Nt = warmup + steps
    c, prob = make_problem(pa, n, W, Nt, steps_per_launch=steps_per_launch, block_size=block_size,
                           amplitudes=amplitudes, **engine_options)
    th_mask = prob.store_time_steps.copy()
    I_mask = np.zeros(Nt, dtype=bool)
    eng = Engine(prob, th_mask, I_mask)
    if prelude is not None:
        prelude()
    eng.advance(warmup)
    torch.cuda.synchronize()
    launches = []
    stream = torch.cuda.current_stream()
    t_start = time.perf_counter()
    end = eng.t + steps
    while eng.t < end:
        k = min(eng.steps_per_launch, end - eng.t)
        if events:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            eng.launch(k)
            e1.record(stream)
            launches.append((k, e0, e1))
        else:
            eng.launch(k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    launch_ms = [(k, e0.elapsed_time(e1)) for k, e0, e1 in launches]
    th, _, nv = eng.results() if download else (None, None, None)
    sp = eng.host.solve_plan
    lat = eng.host.lattice if eng.host.uses_lattice else None
    info = dict(Nj=c.junction_count(), Nf=c.face_count(), bytes_per_step=eng.bytes_per_step(),
                bytes_per_stored_step=8 * c.junction_count() * W,
                stored_in_window=int(th_mask[warmup:].sum()),
                nnz_factor=int(len(sp.lo_val) + len(sp.inv_cm)), nnz_L=int(sp.nnz_L),
                n_stages=sp.n_stages, block_size=eng.block_size, parts=eng.host.parts,
                steps_per_launch=eng.steps_per_launch,
                solver="lattice" if lat is not None else "cholesky",
                lattice=None if lat is None else {"rows": lat.R, "cols": lat.C,
                                                  "table_bytes": lat.table_bytes()})
    eng.close()
    return wall, launch_ms, th, nv, c, info


def factor_path(pa, torch, n, W, steps):
    """The same workload with the face system solved by the staged Cholesky
    factor (the solver for every other circuit): one timed launch of ``steps``
    steps after one untimed step."""
    wall, launch_ms, _, _, _, info = run_gpu(pa, torch, n, W, steps, 1, steps, None, solver="cholesky",
                                             download=False)
    k, ms = launch_ms[0]
    achieved = (info["bytes_per_step"] * k + info["bytes_per_stored_step"] * info["stored_in_window"]) / \
        (ms * 1e-3) / 1e9
    return {"solver": "cholesky", "steps": k, "ms_per_step": ms / k,
            "value": info["Nj"] * W * k / (ms * 1e-3), "unit": "junction-steps/s",
            "bytes_per_step": info["bytes_per_step"], "frac": achieved / HBM_PEAK_GBS,
            "solve_stages": info["n_stages"], "nnz_factor_per_pass": info["nnz_factor"]}


def cpu_baseline(pa, n, W, budget_s):
    """Oracle (numpy/scipy, same calls as the reference) on a bounded sample of steps."""
    from oracle import te_oracle as to
    c, _ = make_problem(pa, n, W, 10)
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.0, 2.0, W)[None, :, None]

    def timed(Nt):
        np.random.seed(0)
        t0 = time.perf_counter()
        to.evolve(c, dt=0.05, Nt=Nt, f=0.25, Is=Is, T=0.01, store_time_steps=[Nt - 1],
                  store_current=False, store_voltage=False)
        return time.perf_counter() - t0
    t_probe = timed(3)
    steps = max(3, int(budget_s / max(t_probe / 3, 1e-6)))
    t = timed(steps)
    rate = c.junction_count() * W * steps / t
    return dict(value=rate, unit="junction-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of SquareArray({n},{n}) W={W} T=0.01 (includes one "
                       f"SuperLU factorisation), single-threaded numpy/scipy; "
                       f"{t:.1f} s total")


def _cpu_slice(args):
    """One worker of the all-cores baseline: the oracle on columns w0..w1 of the
    bench's IV sweep for `steps` steps; returns the evolve call's seconds."""
    n, W, w0, w1, steps = args
    import pyjjasim_amd as pa
    from oracle import te_oracle as to
    c, _ = make_problem(pa, n, W, 10)
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.0, 2.0, W)[None, w0:w1, None]
    np.random.seed(w0)
    t0 = time.perf_counter()
    to.evolve(c, dt=0.05, Nt=steps, f=0.25, Is=Is, T=0.01, store_time_steps=[steps - 1],
              store_current=False, store_voltage=False)
    return time.perf_counter() - t0


def cpu_baseline_all_cores(n, W, steps, max_workers=16):
    """The oracle with the W columns sharded over worker processes (one thread
    each), as the reference's own multiprocessing example shards work; the
    rate uses the slowest worker's time (process start-up excluded)."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    workers = max(1, min(max_workers, len(os.sched_getaffinity(0)), W))
    bounds = np.linspace(0, W, workers + 1).astype(int)
    jobs = [(n, W, int(a), int(b), steps) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    os.environ.update({k: "1" for k in env})
    try:
        with ProcessPoolExecutor(max_workers=len(jobs), mp_context=multiprocessing.get_context("spawn")) as ex:
            times = list(ex.map(_cpu_slice, jobs))
    finally:
        for k, v in env.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    Nj = 2 * n * (n - 1)
    return dict(value=Nj * W * steps / max(times), unit="junction-steps/s", cores=len(jobs), kind="port",
                sample=f"{steps} steps of SquareArray({n},{n}) W={W} T=0.01, columns sharded over "
                       f"{len(jobs)} single-threaded processes; slowest worker {max(times):.1f} s")


def adjoint_bench(pa, torch, n=64, W=32, Nt=4000, checkpoint=200, n_fd=50, fd_steps=200, h=1e-4):
    """Differentiation workload: SquareArray(n, n), trainable per-junction Ic
    (log-normal, sigma 0.2, seed A), loss = mean over W bias points (0.5..2.0 Ic)
    of (V_w - V*_w)^2 with V_w the time-averaged voltage of the biased junctions
    and V* from a hidden Ic field (seed B); beta_C = 1, f = 0.1, T = 0, dt = 0.05.

    Timed: the forward pass and the backward pass (segment recompute plus the
    adjoint sweep), each after a warm-up. Checked: the gradient against central
    differences on ``n_fd`` seeded junctions over the first ``fd_steps`` steps,
    Richardson-extrapolated from steps h and h/2 (truncation error O(h^4));
    over the full horizon the dynamics are chaotic and central differences do
    not converge as h shrinks, so no difference quotient can check it there."""
    c = pa.SquareArray(n, n)
    c.set_capacitance(1.0)
    Nj, Nf = c.junction_count(), c.face_count()
    dt = 0.05
    base = c.current_base(0.0)
    biased = torch.tensor(np.abs(base) > 0, device="cuda")
    dev = torch.device("cuda")
    Is = torch.tensor(base[:, None] * np.linspace(0.5, 2.0, W)[None, :], dtype=torch.float64, device=dev)
    f = torch.full((Nf,), 0.1, dtype=torch.float64, device=dev)
    Ic0 = torch.tensor(np.exp(0.2 * np.random.default_rng(101).standard_normal(Nj)), dtype=torch.float64, device=dev)
    Ic_star = torch.tensor(np.exp(0.2 * np.random.default_rng(202).standard_normal(Nj)), dtype=torch.float64, device=dev)

    def make(steps):
        store = np.zeros(steps, dtype=bool)
        store[-1] = True
        evo = pa.DifferentiableEvolution(c, time_step=dt, time_step_count=steps, problem_count=W,
                                         store_time_steps=store, checkpoint_every=checkpoint)

        def volt(Ic):
            th = evo(Ic, f, Is)[:, :, -1]
            return th[biased].mean(0) / (steps * dt)
        with torch.no_grad():
            target = volt(Ic_star)
        return (lambda Ic: ((volt(Ic) - target) ** 2).mean()), evo

    loss, evo = make(Nt)
    times = []
    try:
        for rep in range(2):                          # the first pass warms up
            Ic = Ic0.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L = loss(Ic)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            L.backward()
            torch.cuda.synchronize()
            times.append((t1 - t0, time.perf_counter() - t1))
        parts = evo.eng.host.parts
    finally:
        evo.close()
    t_fwd, t_bwd = times[-1]
    loss_fd, evo_fd = make(fd_steps)
    try:
        Ic = Ic0.clone().requires_grad_(True)
        loss_fd(Ic).backward()
        g = Ic.grad.detach().cpu().numpy()
        idx = np.random.default_rng(303).choice(Nj, size=n_fd, replace=False)

        def central(j, step):
            p, m = Ic0.clone(), Ic0.clone()
            p[j] += step
            m[j] -= step
            with torch.no_grad():
                return (loss_fd(p) - loss_fd(m)).item() / (2 * step)
        fd = np.array([(4 * central(j, h / 2) - central(j, h)) / 3 for j in idx])
    finally:
        evo_fd.close()
    gs = g[idx]
    scale = np.abs(g).max()
    return {"workload": f"SquareArray({n},{n}) W={W} Nt={Nt} checkpoint={checkpoint}, d loss / d Ic",
            "parts": int(parts), "forward_s": t_fwd, "backward_s": t_bwd, "grad_over_fwd": t_bwd / t_fwd,
            "fd_check": {"steps": fd_steps, "junctions": int(n_fd), "h": h, "richardson": "h, h/2",
                         "max_abs_err_over_max_grad": float(np.abs(gs - fd).max() / scale),
                         "max_rel_err": float(np.max(np.abs(gs - fd) / np.abs(fd))),
                         "max_abs_grad": float(scale)}}


def uneven_problems(pa, Nt, n_problems=64, seed=3003):
    """Uneven batch: n_problems seeded draws of lattice (square, honeycomb,
    triangular) and count n in [16, 256]; junction self-inductance L = 1, R = 1,
    C = 0; f ~ U[0, 0.5]; Shapiro drive Is(t) = (Idc + 0.5 sin(0.3 t)) *
    current_base(0) with Idc ~ U[0.5, 1.5]; T = 0, dt = 0.05; W = 1 each; theta
    stored at the last step."""
    rng = np.random.default_rng(seed)
    kinds = [pa.SquareArray, pa.HoneycombArray, pa.TriangularArray]
    dt = 0.05
    wave = np.sin(0.3 * dt * np.arange(Nt))
    probs, names = [], []
    for _ in range(n_problems):
        K = kinds[rng.integers(3)]
        n = int(rng.integers(16, 257))
        f = float(rng.uniform(0.0, 0.5))
        Idc = float(rng.uniform(0.5, 1.5))
        c = K(n, n)
        c.set_inductance(1.0)
        drive = pa.SeparableSource(c.current_base(0.0), Idc + 0.5 * wave)
        probs.append(pa.TimeEvolutionProblem(c, time_step=dt, time_step_count=Nt, external_flux=f,
                                             current_sources=drive, store_time_steps=[Nt - 1],
                                             store_voltage=False, store_current=False))
        names.append(f"{K.__name__}({n},{n})")
    return probs, names


def uneven_bench(pa, torch, steps, warmup, plain_steps, n_problems=64):
    """The uneven batch in one launch per chunk, autotuned; beside it the same
    problems one at a time (the plain loop) on a sample of steps."""
    from pyjjasim_amd.engine import BatchEngine
    Nt = warmup + steps
    t0 = time.perf_counter()
    probs, names = uneven_problems(pa, Nt, n_problems)
    tb = pa.TimeEvolutionBatch(probs, autotune=True)
    eng = tb.make_engine()
    setup_s = time.perf_counter() - t0
    Njs = np.array([p.circuit._Nj() for p in probs])
    try:
        eng.advance(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.advance(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        bytes_per_step = eng.bytes_per_step()
        for p, (th, _, nv) in zip(probs, eng.results()):
            A = p.circuit.get_cycle_matrix()
            n_host = (-A @ np.round(th[:, :, -1] / (2.0 * np.pi))).astype(int)
            assert np.array_equal(n_host, nv[:, :, -1]), "device vortex counts disagree with theta"
        hosts = eng.hosts
    finally:
        eng.close()
    value = float(Njs.sum()) * steps / wall
    plain = None
    plain_steps = min(plain_steps, Nt - 5)   # each plain run starts with 5 untimed steps
    if plain_steps > 0:
        per_step = 0.0
        for i, p in enumerate(probs):
            e = BatchEngine([p], [(np.zeros(Nt, bool), np.zeros(Nt, bool))], hosts=[hosts[i]])
            try:
                e.launch(5)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                e.advance(plain_steps)
                torch.cuda.synchronize()
                per_step += (time.perf_counter() - t1) / plain_steps
            finally:
                e.close()
        plain = {"value": float(Njs.sum()) / per_step, "unit": "junction-steps/s",
                 "note": f"the same problems one at a time, {plain_steps} steps each",
                 "batched_over_plain": value * per_step / float(Njs.sum())}
    big = int(np.argmax(Njs))
    return {"metric": "junction-steps/s (uneven batch, FP64)", "value": value,
            "unit": "junction-steps/s", "steps": steps, "warmup": warmup,
            "ms_per_step": wall / steps * 1e3, "dtype": "f64",
            "config": {"workload": f"uneven_batch_{n_problems}", "problems": n_problems,
                       "junctions_total": int(Njs.sum()), "junctions_min": int(Njs.min()),
                       "junctions_max": int(Njs.max()), "largest": names[big],
                       "global_store_problems": int(sum(h.global_store for h in hosts)),
                       "tuning": {k: (v if k != "s_per_step" else {c: t * 1e3 for c, t in v.items()})
                                  for k, v in tb.tuning.items()},
                       "tuning_unit": "ms per step, by workgroup size x parts per column"},
            "hbm_GBps_algorithmic": bytes_per_step * steps / wall / 1e9,
            "host_setup_s": setup_s, "plain": plain}


def large_bench(pa, torch, steps, checkpoint, n=512, W=16, full_steps=50000, solver="auto"):
    """Large workload: SquareArray(n, n), W bias points linspace(0.2, 1.2) *
    current_base(0), beta_C = 1, junction self-inductance beta_L = 0.5, per-face
    flux f = 0.1 + 0.02 z (seeded), T = 0.05, dt = 0.05; the forward pass and the
    adjoint gradient d loss / d f (loss: mean time-averaged voltage of the biased
    junctions), checkpoints every ``checkpoint`` steps. Timed over ``steps``
    steps (a bounded sample of the full horizon; every step costs the same),
    after one untimed pass."""
    dt = 0.05
    t0 = time.perf_counter()
    c = pa.SquareArray(n, n)
    c.set_capacitance(1.0)
    c.set_inductance(0.5)
    Nj, Nf = c.junction_count(), c.face_count()
    base = c.current_base(0.0)
    dev = torch.device("cuda")
    biased = torch.tensor(np.abs(base) > 0, device=dev)
    Is = torch.tensor(base[:, None] * np.linspace(0.2, 1.2, W)[None, :], dtype=torch.float64, device=dev)
    f0 = 0.1 + 0.02 * np.random.default_rng(505).standard_normal(Nf)
    Ic = torch.ones(Nj, dtype=torch.float64, device=dev)
    store = np.zeros(steps, dtype=bool)
    store[-1] = True
    evo = pa.DifferentiableEvolution(c, time_step=dt, time_step_count=steps, problem_count=W,
                                     temperature=0.05, store_time_steps=store,
                                     checkpoint_every=min(checkpoint, steps), seed=7, solver=solver)
    setup_s = time.perf_counter() - t0
    for rep in range(2):                          # the first pass warms up
        f = torch.tensor(f0, dtype=torch.float64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        th = evo(Ic, f, Is)[:, :, -1]
        loss = th[biased].mean() / (steps * dt)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    g = f.grad.detach()
    fwd, bwd = (t1 - t0) / steps, (t2 - t1) / steps
    face_store = "global" if evo.eng.host.global_store else "lds"
    n_exec = int(evo.eng.host.n_solve_exec)
    used = "lattice" if evo.eng.host.uses_lattice else "cholesky"
    parts = int(evo.eng.host.parts)
    # forward pass against the HBM roofline: the step's algorithmic bytes
    # (W (24 Nj + 48 Nf) + solve tables or factor + 40 Nj, plus the dense bias
    # Is, 8 Nj W, passed as a tensor and read every step) over the measured
    # time per step (the launches include the checkpoint copies)
    bps = int(evo.eng.bytes_per_step()) + 8 * Nj * W
    achieved = bps / fwd / 1e9
    evo.close()
    return {"metric": "junction-steps/s (large array forward, FP64)", "value": Nj * W / fwd,
            "unit": "junction-steps/s", "steps": steps, "dtype": "f64",
            "config": {"workload": f"large_square{n}_W{W}", "junctions": Nj, "faces": Nf, "W": W,
                       "beta_L": 0.5, "T": 0.05, "checkpoint": min(checkpoint, steps),
                       "face_store": face_store, "solver": used, "parts": parts,
                       "solve_records_per_column_step": n_exec},
            "forward_ms_per_step": fwd * 1e3, "backward_ms_per_step": bwd * 1e3,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_step": bps, "pass": "forward"},
            "grad_over_fwd": bwd / fwd,
            "full_horizon_estimate_s": {"steps": full_steps, "forward": fwd * full_steps,
                                        "backward": bwd * full_steps},
            "grad_f_finite": bool(torch.isfinite(g).all().item()),
            "host_setup_s": setup_s}


TRAFFIC_PROFILE = os.path.join("profiles", "evolve_kernel_square100_W256_traffic.txt")


def measured_traffic():
    """HBM bytes per step of evolve_kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE of one dispatch, each scaled by the
    copy-kernel factor, divided by the dispatch's steps), the launch length
    and the face-system solver it was taken with; None when the summary is
    missing."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if ":" in line and not line.startswith(("#", " ", "\t")):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_step"]), int(vals["steps_per_launch"]), vals.get("solver", "cholesky")
    except (KeyError, ValueError):
        return None


def host_info():
    """The host the CPU baseline ran on: CPU model, usable cores, library
    versions and thread settings."""
    import scipy
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    return {"cpu_model": model, "usable_cores": len(os.sched_getaffinity(0)), "numpy": np.__version__,
            "scipy": scipy.__version__, "thread_env": env}


def plain_best(pa, torch, n, steps, amplitude=1.0, parts_options=(1, 2, 4, 8, 16)):
    """The plain loop: one problem (one bias point, W = 1) at a time, in its
    fastest single-problem layout: one workgroup with the LDS face store (the
    lattice solve), or the column split over 2..16 workgroups with the global
    face store (the factor). Returns the best rate and every layout's ms/step."""
    per = {}
    for parts in parts_options:
        opts = {"parts": parts} if parts > 1 else {"parts": 1}
        if parts > 1:
            opts["face_store"] = "global"
        wall, _, _, _, c, info = run_gpu(pa, torch, n, 1, steps, 100, None, None, events=False,
                                         amplitudes=[amplitude], **opts)
        per[f"parts{parts}_{info['solver']}"] = wall / steps * 1e3
    best = min(per, key=per.get)
    Nj = 2 * n * (n - 1)
    return {"value": Nj / (per[best] * 1e-3), "unit": "junction-steps/s", "layout": best,
            "ms_per_step_by_layout": per,
            "note": f"one problem (W=1, bias {amplitude} current_base) per run, {steps} steps, "
                    "fastest single-problem layout"}


def plain_small(pa, torch, Nt=10000):
    """The reference's plain small case: SquareArray(10,10), Ic = R = 1, C = L = 0,
    f = 0.1, Is = 0.5 current_base(0), T = 0, dt = 0.05, Nt steps, W = 1, every
    step stored (theta, current, voltage); the whole compute() call is timed,
    download of the stored steps included."""
    c = pa.SquareArray(10, 10)
    Is = 0.5 * c.current_base(0.0)[:, None, None]
    prob = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=Nt, external_flux=0.1,
                                   current_sources=Is)
    prob.compute()                                  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = prob.compute()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    Nj = c.junction_count()
    assert res.get_theta().shape == (Nj, 1, Nt)
    return {"workload": "plain_small_square10_W1", "steps": Nt, "seconds": wall,
            "value": Nj * Nt / wall, "unit": "junction-steps/s",
            "note": "compute() end to end: plan, launches, download of every stored step"}


def copy_calibration(torch, mib):
    """One plain device copy of `mib` MiB: the reference point for the HBM
    counters (FETCH_SIZE and WRITE_SIZE of this kernel should read `mib` MiB)."""
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--n", type=int, default=100)
    ap.add_argument("--W", type=int, default=256)
    ap.add_argument("--steps-per-launch", type=int, default=None)
    ap.add_argument("--block-size", type=int, default=None)
    ap.add_argument("--plain-steps", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-plain-small", action="store_true")
    ap.add_argument("--solver", choices=["auto", "lattice", "cholesky"], default="auto",
                    help="face-system solver of the timed run")
    ap.add_argument("--factor-steps", type=int, default=2000,
                    help="steps of the same workload with the factor solver, beside a lattice-solve run (0: skip)")
    ap.add_argument("--workload", choices=["iv_sweep", "uneven_batch", "large"], default="iv_sweep",
                    help="iv_sweep: SquareArray(100,100) W=256 (default); uneven_batch: "
                         "64 problems of uneven size in one launch per chunk; large: "
                         "SquareArray(512,512) W=16 forward + adjoint on --steps steps")
    ap.add_argument("--checkpoint", type=int, default=1000, help="large: checkpoint interval")
    ap.add_argument("--problems", type=int, default=64, help="uneven_batch: problem count")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(columns and problems are independent and would shard over GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import pyjjasim_amd as pa
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration(torch, args.copy_calibration)
        return
    if args.workload == "large":
        out = large_bench(pa, torch, args.steps, args.checkpoint, solver=args.solver)
        out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                    "data": "synthetic (seeded flux disorder, seeded in-kernel thermal noise)"})
        print(json.dumps(out), flush=True)
        return
    if args.workload == "uneven_batch":
        out = uneven_bench(pa, torch, args.steps, args.warmup, args.plain_steps, args.problems)
        out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                    "data": "synthetic (seeded lattices, sizes, flux and drive)"})
        print(json.dumps(out), flush=True)
        return

    # the plain loop (one problem at a time) and the factor path are measured
    # between the batched run's host set-up and its warmup steps; their GPU work
    # also brings the device to its working clock right before the timed steps
    plain = factor = None

    def prelude():
        # the plain loop, then the same batch on the factor path (full width, so
        # the device is at its loaded clock when the timed steps start)
        nonlocal plain, factor
        if args.plain_steps > 0:
            plain = plain_best(pa, torch, args.n, args.plain_steps)
        if args.solver != "cholesky" and args.factor_steps > 0:
            factor = factor_path(pa, torch, args.n, args.W, args.factor_steps)

    wall, launch_ms, th, nv, c, info = run_gpu(pa, torch, args.n, args.W, args.steps, args.warmup,
                                               args.steps_per_launch, args.block_size,
                                               prelude=prelude, solver=args.solver)
    Nj, W, K = info["Nj"], args.W, args.steps
    value = Nj * W * K / wall
    # dominant kernel: evolve_kernel; one launch advances k steps of the whole
    # batch. Algorithmic bytes of the timed launches: B_step per step plus the
    # theta rows stored within them
    tot_steps = sum(k for k, _ in launch_ms)
    tot_ms = sum(ms for _, ms in launch_ms)
    n_launch = len(launch_ms)
    mean_launch_ms = tot_ms / n_launch
    bytes_timed = info["bytes_per_step"] * tot_steps + info["bytes_per_stored_step"] * info["stored_in_window"]
    bytes_per_launch = bytes_timed / n_launch
    achieved = bytes_per_launch / (mean_launch_ms * 1e-3) / 1e9
    traffic = traffic_note = None
    if (args.n, W) == (100, 256):
        mt = measured_traffic()
        if mt is not None and mt[2] == info["solver"]:
            per_step, prof_len, _ = mt
            traffic = per_step * tot_steps / n_launch
            traffic_note = (f"{TRAFFIC_PROFILE}: {per_step / 1e6:.1f} MB/step measured on a "
                            f"{prof_len}-step dispatch, times {tot_steps / n_launch:g} steps per launch")
    # integrity check at full size: device vortex counts equal the host count
    A = c.get_cycle_matrix()
    n_host = (-A @ np.round(th[:, :, -1] / (2.0 * np.pi))).astype(int)
    assert np.array_equal(n_host, nv[:, :, -1]), "device vortex counts disagree with theta"
    if plain is not None:
        plain["batched_over_plain"] = value / plain["value"]
    out = {
        "metric": "junction-steps/s (batched IV sweep, FP64)",
        "value": value,
        "unit": "junction-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Ic disorder, seeded in-kernel thermal noise)",
        "config": {"workload": WORKLOAD, "lattice": f"SquareArray({args.n},{args.n})",
                   "junctions": Nj, "faces": info["Nf"], "W": W, "dt": 0.05, "T": 0.01,
                   "beta_C": 1.0, "f": 0.25, "Nt": args.warmup + K,
                   "bias": "SeparableSource(current_base(0), linspace(0, 2, W))",
                   "steps_per_launch": info["steps_per_launch"], "block_size": info["block_size"],
                   "parts": info["parts"], "solver": info["solver"],
                   "solve_stages": info["n_stages"] if info["solver"] == "cholesky" else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": "evolve_kernel", "bytes_per_step": info["bytes_per_step"],
                     "bytes_per_launch": bytes_per_launch, "stored_steps_in_window": info["stored_in_window"],
                     "mean_launch_ms": mean_launch_ms, "launches": n_launch},
        "plain": plain,
    }
    if info["solver"] == "cholesky":
        out["roofline"].update({"nnz_factor_per_pass": info["nnz_factor"], "nnz_L": info["nnz_L"]})
    else:
        out["roofline"]["lattice"] = info["lattice"]
    if info["solver"] == "lattice" and factor is not None:
        out["factor_path"] = factor
    if not args.no_plain_small:
        out["plain_small"] = plain_small(pa, torch)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(pa, torch)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pa, args.n, W, args.cpu_seconds)
        out["cpu_baseline"].update(host_info())
        steps = int(out["cpu_baseline"]["sample"].split()[0])
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(args.n, W, steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
