"""Time evolution as a differentiable function of the critical currents, the
external flux and the bias currents (stencil 3, DefaultCPR).

``DifferentiableEvolution(circuit, ...)`` factorises the face matrix once (it
does not depend on Ic, f or Is) and is then called with torch tensors:

    evo = DifferentiableEvolution(circuit, time_step=0.05, time_step_count=4000,
                                  problem_count=32, store_time_steps=[3999])
    theta = evo(Ic, f, Is)          # (Nj, W, n_stored), differentiable
    loss(theta).backward()          # Ic.grad, f.grad, Is.grad

Forward: the GPU step kernel (the same as TimeEvolutionProblem.compute()) with
the phases saved every ``checkpoint_every`` steps. Backward: for each segment,
last first, the phases of the segment are recomputed from its checkpoint
(Philox noise is counter-based, so a recomputed step draws the same deviates)
and the adjoint kernel (include/jja.h, jja_adjoint) runs the segment's steps in
reverse. The step's linear part M = D A^T K^-1 A D - D is symmetric, so the
backward solve uses the forward's factor and solve program unchanged.

Reference scheme: time_evolution.py:523-580 (the step), :547-558 (x), :560-570
(the solve); the reference itself has no gradients, so the adjoint has no
counterpart there.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .engine import Engine, _torch
from .time_evolution import TimeEvolutionProblem

__all__ = ["DifferentiableEvolution"]


def _term(tensor, n, W, name):
    """Device layout of an (n,) or (n, W) constant input and its jja_source."""
    t = tensor.detach()
    src = _native.Source()
    if t.dim() == 1:
        if t.shape[0] != n:
            raise ValueError(f"{name} must have shape ({n},) or ({n}, {W})")
        dev = t.contiguous()
        src.s_i, src.s_w, src.s_t = 1, 0, 0
    elif t.dim() == 2 and tuple(t.shape) == (n, W):
        dev = t.t().contiguous()
        src.s_i, src.s_w, src.s_t = 1, n, 0
    else:
        raise ValueError(f"{name} must have shape ({n},) or ({n}, {W})")
    src.base = dev.data_ptr()
    src.t0 = 0
    src.zero = 0
    return dev, src


class DifferentiableEvolution:
    """Differentiable stencil-3 time evolution of one circuit for W columns.

    Arguments follow TimeEvolutionProblem (time_evolution.py:94-103); the
    external flux and the bias currents are the call's tensors and are constant
    in time. ``checkpoint_every`` steps are recomputed per backward segment.
    """

    def __init__(self, circuit, time_step=0.05, time_step_count=1000, problem_count=1,
                 temperature=0.0, store_time_steps=None, config_at_minus_1=None,
                 config_at_minus_2=None, seed=None, checkpoint_every=200, block_size=None,
                 face_store="auto", parts=None, local_stages="auto", column_pairs=False, solver="auto"):
        torch = self.torch = _torch()
        W, Nj, Nf = int(problem_count), circuit._Nj(), circuit._Nf()
        self.W, self.Nj, self.Nf, self.Nt = W, Nj, Nf, int(time_step_count)
        self.C = int(checkpoint_every)
        if self.C < 1:
            raise ValueError("checkpoint_every must be at least 1")
        # placeholders: the flux and bias of each call replace these sources
        prob = TimeEvolutionProblem(
            circuit, time_step=time_step, time_step_count=time_step_count,
            external_flux=np.zeros((1, W, 1)), current_sources=np.zeros((1, W, 1)),
            temperature=temperature, store_time_steps=store_time_steps, store_theta=True,
            store_voltage=False, store_current=False, config_at_minus_1=config_at_minus_1,
            config_at_minus_2=config_at_minus_2, seed=seed if seed is not None else 0,
            steps_per_launch=self.C, block_size=block_size, face_store=face_store, parts=parts,
            local_stages=local_stages, column_pairs=column_pairs, solver=solver,
            max_resident_output_bytes=1 << 62)   # stored phases stay on the device
        self.problem = prob
        self.th_mask = prob.store_time_steps.copy()
        self.eng = Engine(prob, self.th_mask, np.zeros(self.Nt, dtype=bool))
        self.n_th = self.eng.n_th
        self.theta0 = self.eng.theta.clone()
        pos = self.eng.host.face_of_pos.astype(np.int64)
        self.face_of_pos = torch.from_numpy(pos).to(self.eng.dev)
        self.n_seg = -(-self.Nt // self.C)

    def __call__(self, Ic, f, Is):
        return _apply(Ic, f, Is, self)

    def close(self):
        """Free the device plan now (also done when the object is collected)."""
        eng = getattr(self, "eng", None)
        if eng is not None:
            eng.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def _check_status(self):
        """Raise NativeError if a column barrier of a split launch gave up (its
        results are invalid); waits for the launches of a split plan."""
        eng = self.eng
        if eng.host.parts > 1:
            self.torch.cuda.current_stream().synchronize()
        _native.check(eng.lib.jja_plan_status(eng.handle))

    # ------------------------------------------------------------------ forward
    def _set_inputs(self, Ic, f, Is):
        eng = self.eng
        self._Ic = Ic.detach().contiguous()
        if self._Ic.shape != (self.Nj,):
            raise ValueError(f"Ic must have shape ({self.Nj},)")
        self._f, self._fsrc = _term(f, self.Nf, self.W, "f")
        self._Is, self._isrc = _term(Is, self.Nj, self.W, "Is")
        eng.srcs["f"].src = self._fsrc
        eng.srcs["Is"].src = self._isrc
        eng.run.Ic = self._Ic.data_ptr()

    def forward(self, Ic, f, Is):
        """Stored theta (Nj, W, n_stored) and the state the backward needs."""
        eng = self.eng
        self._set_inputs(Ic, f, Is)
        eng.theta.copy_(self.theta0)
        eng.t = 0
        ckpt = []
        for seg in range(self.n_seg):
            ckpt.append(eng.theta.clone())
            eng.launch(min(self.C, self.Nt - eng.t))
        self._check_status()
        out = eng.th_out[:self.n_th].permute(2, 1, 0).clone()
        return out, (ckpt, self._Ic, self._f, self._Is, self._fsrc, self._isrc)

    # ----------------------------------------------------------------- backward
    def backward(self, state, g):
        """Gradients (Ic (Nj,), f (W, Nf), Is (W, Nj)) of <g, theta_stored>."""
        torch, eng = self.torch, self.eng
        ckpt, Ic, f, Is, fsrc, isrc = state
        # the inputs of this graph (another call may have run since)
        self._Ic, self._f, self._Is = Ic, f, Is
        eng.srcs["f"].src, eng.srcs["Is"].src = fsrc, isrc
        eng.run.Ic = Ic.data_ptr()
        W, Nj, Nf, Nt, C_ = self.W, self.Nj, self.Nf, self.Nt, self.C
        dev, f64 = eng.dev, torch.float64
        stream = torch.cuda.current_stream().cuda_stream
        g_dev = g.detach().permute(2, 1, 0).contiguous() if self.n_th else None
        lam = torch.zeros((2, W, Nj), dtype=f64, device=dev)
        u = torch.zeros((W, Nj), dtype=f64, device=dev)
        gIc = torch.zeros((W, Nj), dtype=f64, device=dev)
        gIs = torch.zeros((W, Nj), dtype=f64, device=dev)
        gf = torch.zeros((W, Nf), dtype=f64, device=dev)
        hist = torch.empty((C_ + 1, W, Nj), dtype=f64, device=dev)
        run = eng.run
        saved = (run.th_slot, run.slot_t0, run.th_out)
        adj = _native.AdjointRun()
        adj.W = W
        adj.Ic = self._Ic.data_ptr()
        adj.hist = hist.data_ptr()
        adj.lam = lam.data_ptr()
        adj.u = u.data_ptr()
        if g_dev is not None:
            adj.g = g_dev.data_ptr()
            adj.g_slot = eng.th_slot.data_ptr()
            adj.slot_t0, adj.g_lo, adj.g_hi = 0, 0, Nt
        adj.grad_Ic, adj.grad_Is, adj.grad_f = gIc.data_ptr(), gIs.data_ptr(), gf.data_ptr()
        adj.zbuf = eng.zbuf.data_ptr() if eng.zbuf is not None else None
        try:
            for seg in reversed(range(self.n_seg)):
                s0, s1 = seg * C_, min((seg + 1) * C_, Nt)
                ck = ckpt[seg]
                eng.theta.copy_(ck)
                hist[0].copy_(ck[(s0 - 2) % 2])
                hist[1].copy_(ck[(s0 - 1) % 2])
                if s1 - s0 > 1:
                    # recompute theta_m for m in [s0, s1-2] into hist[m - s0 + 2]
                    slots = torch.arange(2, s1 - s0 + 1, dtype=torch.int32, device=dev)
                    run.th_slot, run.slot_t0, run.th_out = slots.data_ptr(), s0, hist.data_ptr()
                    eng.t = s0
                    eng.launch(s1 - s0 - 1)
                    run.th_slot, run.slot_t0, run.th_out = saved
                adj.prime = 1 if seg == self.n_seg - 1 else 0
                adj.hist_m0 = s0 - 2
                _native.check(eng.lib.jja_adjoint(eng.handle, C.byref(adj), s1 - 1, s0, stream))
            self._check_status()
        finally:
            run.th_slot, run.slot_t0, run.th_out = saved
        gf_face = torch.zeros_like(gf)
        gf_face[:, self.face_of_pos] = gf
        return gIc.sum(0), gf_face, gIs


def _make_fn():
    torch = _torch()

    class EvolveFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, Ic, f, Is, evo):
            ctx.evo = evo
            ctx.f_shape, ctx.Is_shape = tuple(f.shape), tuple(Is.shape)
            out, ctx.state = evo.forward(Ic, f, Is)
            return out

        @staticmethod
        def backward(ctx, g):
            gIc, gf, gIs = ctx.evo.backward(ctx.state, g)
            gf = gf.sum(0) if len(ctx.f_shape) == 1 else gf.t()
            gIs = gIs.sum(0) if len(ctx.Is_shape) == 1 else gIs.t()
            return gIc, gf, gIs, None
    return EvolveFn


_FN = []   # the autograd Function, made on first use (torch imports lazily)


def _apply(Ic, f, Is, evo):
    if not _FN:
        _FN.append(_make_fn())
    return _FN[0].apply(Ic, f, Is, evo)
