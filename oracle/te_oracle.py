"""CPU oracle for the time-evolution hot path (FP64 numpy/scipy restatement).

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module. The product path
(``pyjjasim_amd``) never calls it.

It restates the reference scheme step for step, with the same library calls and
the same order of floating-point operations:

* ``evolve_core``  <- time_evolution_core (time_evolution.py:461-582)
    setup:   Rv = 1/(dt R), Cv = C/dt^2, c_k = C1[k] Rv + C2[k] Cv  (:470-499)
             K = A (L + diag(1/(Rv+Cv))) A^T, factorised once      (:504-506)
             zero flags via np.allclose for non-time-dependent inputs (:509-519)
    step i:  eta = sqrt(2 T Rv) Z                                  (:533-538)
             X   = Ic sin(extrapolated theta) + sum_k c_k theta_k   (:547-553)
             x   = eta - Is + X                                     (:555-558)
             y   = A^T K^-1 (A (x/c0 - theta_s) - 2 pi f)           (:560-569)
             theta_n = (y - x)/c0 ; I = y + Is                      (:570-577)
             theta_s += Vs dt                                       (:579-580)
* ``evolve``       <- time_evolution (time_evolution.py:422-458): widened store
                      masks, stencil voltage, L dI/dt, prefix removal.
* ``vortex``       <- TimeEvolutionResult.get_vortex_configuration (:734-755).
* ``default_noise`` <- the legacy numpy RNG call order of :533-537 (fresh
                      randn each step when Nj <= 500, otherwise every third step
                      with row permutations in between).

Parity pins (tests/test_oracle.py): the reference's own trajectories written by
tests/golden/make_golden.py (T=0 and seeded T>0, all stencils, inductance,
voltage sources, time-dependent and callable inputs); invariants (fluxoid
constraint, current conservation, telescoping average voltage, zero-input fixed
point); a dense brute-force step on tiny circuits.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse
import scipy.sparse.linalg

STENCILS = {
    3: ((1.0, -1.0, 0.0), (1.0, -2.0, 1.0)),
    4: ((1.5, -2.0, 0.5, 0.0), (2.0, -5.0, 4.0, -1.0)),
    5: ((11.0 / 6, -3.0, 1.5, -1.0 / 3, 0.0), (35.0 / 12, -26.0 / 3, 9.5, -14.0 / 3, 11.0 / 12)),
}


def _is_timedep(x):
    if callable(x):
        return True
    s = np.shape(x)
    return len(s) > 0 and s[-1] > 1


def _problem_count(*xs):
    def cnt(x):
        s = np.shape(x(0) if callable(x) else x)
        return s[1] if len(s) > 1 else 1
    return max(cnt(x) for x in xs)


class Inputs:
    """Per-step access to f (Nf,W), Is/Vs/T (Nj,W) with the reference's broadcasting
    (time_evolution.py:119-130, 342-356)."""

    def __init__(self, Nj, Nf, Nt, f, Is, Vs, T):
        self.W = _problem_count(f, Is, Vs, T)
        self.Nj, self.Nf, self.Nt = Nj, Nf, Nt
        shapes = {"f": Nf, "Is": Nj, "Vs": Nj, "T": Nj}
        self.timedep = {}
        self.src = {}
        for name, x in (("f", f), ("Is", Is), ("Vs", Vs), ("T", T)):
            self.timedep[name] = _is_timedep(x)
            self.src[name] = x if callable(x) else np.broadcast_to(np.array(x), (shapes[name], self.W, Nt))

    def get(self, name, i):
        x = self.src[name]
        n = self.Nf if name == "f" else self.Nj
        if callable(x):
            return np.broadcast_to(x(i), (n, self.W))
        return x[:, :, i]


def default_noise(Nj, W):
    """Noise source reproducing the reference's legacy-RNG call sequence."""
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


def philox_noise(seed, W, Nj, column_offset=0):
    from .philox import normals
    cols = np.arange(W) + column_offset
    return lambda i: normals(seed, i, cols, Nj)


def superlu_solver(K):
    return scipy.sparse.linalg.factorized(K)


def evolve_core(circuit, dt, Nt, inputs: Inputs, th_store_mask, I_store_mask, configs,
                stencil_width=3, cp=None, noise=None, solver=superlu_solver):
    """Returns (th_out, I_out) of shape (Nj, W, prefix + stored), prefix = stencil-1.

    configs: list of stencil_width-1 initial (Nj, W) arrays [theta_-1, theta_-2, ...].
    cp: callable (Ic (Nj,1), theta (Nj,W)) -> current; default Ic*sin(theta).
    noise: callable i -> (Nj, W) standard normals; default the reference's RNG order.
    """
    if cp is None:
        def cp(Ic, th):
            return Ic * np.sin(th)
    Nj, W = circuit._Nj(), inputs.W
    A = circuit.get_cycle_matrix()
    AT = A.T
    Ic = circuit._Ic()[:, None]
    Rv = 1 / (dt * circuit._R()[:, None])
    Cv = circuit._C()[:, None] / (dt ** 2)
    S1, S2 = STENCILS[stencil_width]
    Cnext = Cv + Rv
    c = [S1[k] * Rv + S2[k] * Cv for k in range(stencil_width)]
    sw = stencil_width
    n_prefix = sw - 1
    th_out = np.zeros((Nj, W, int(np.sum(th_store_mask)) + n_prefix))
    I_out = np.zeros((Nj, W, int(np.sum(I_store_mask)) + n_prefix))
    hist = [np.array(cfg, dtype=np.double).reshape(Nj, W).copy() for cfg in configs]
    for k in range(n_prefix):          # hist[0] = theta_-1 sits at prefix position sw-2
        th_out[:, :, sw - 2 - k] = hist[k]
        I_out[:, :, sw - 2 - k] = cp(Ic, hist[k])
    L = circuit._L()
    K = A @ (L + scipy.sparse.diags(1.0 / Cnext[:, 0], 0)) @ AT
    Ksolve = solver(K)
    theta_s = np.zeros((Nj, W))
    Is, T, Vs, f = (inputs.get(n, 0) for n in ("Is", "T", "Vs", "f"))
    zero = {n: (not inputs.timedep[n]) and bool(np.allclose(v, 0))
            for n, v in (("Is", Is), ("T", T), ("Vs", Vs), ("f", f))}
    if noise is None:
        noise = default_noise(Nj, W)
    eta = 0.0
    i_th = i_I = 0
    for i in range(Nt):
        if inputs.timedep["T"]:
            T = inputs.get("T", i)
        if inputs.timedep["Is"]:
            Is = inputs.get("Is", i)
        if inputs.timedep["Vs"]:
            Vs = inputs.get("Vs", i)
        if inputs.timedep["f"]:
            f = inputs.get("f", i)
        if not zero["T"]:
            eta = ((2.0 * T * Rv) ** 0.5) * noise(i)
        t1 = hist[0]
        t2 = hist[1]
        if sw == 3:
            X = cp(Ic, 2 * t1 - t2) + c[1] * t1 + c[2] * t2
        elif sw == 4:
            t3 = hist[2]
            X = cp(Ic, 3 * t1 - 3 * t2 + t3) + c[1] * t1 + c[2] * t2 + c[3] * t3
        else:
            t3, t4 = hist[2], hist[3]
            X = cp(Ic, 4 * t1 - 6 * t2 + 4 * t3 - t4) + c[1] * t1 + c[2] * t2 + c[3] * t3 + c[4] * t4
        x = (eta + X) if zero["Is"] else (eta - Is + X)
        u = x / c[0]
        if not zero["Vs"]:
            u = u - theta_s
        rhs = A @ u
        if not zero["f"]:
            rhs = rhs - 2 * np.pi * f
        y = AT @ Ksolve(rhs)
        th_new = (y - x) / c[0]
        hist = [th_new] + hist[:-1]
        if th_store_mask[i]:
            th_out[:, :, i_th + n_prefix] = th_new
            i_th += 1
        if I_store_mask[i]:
            I_out[:, :, i_I + n_prefix] = y + Is
            i_I += 1
        if not zero["Vs"]:
            theta_s += Vs * dt
    return th_out, I_out


def _derivative(x, idx, stencil, dt):
    w = len(stencil)
    acc = stencil[0] * x[:, :, idx] + stencil[1] * x[:, :, idx - 1]
    if w >= 4:
        acc = acc + stencil[2] * x[:, :, idx - 2]
    if w >= 5:
        acc = acc + stencil[3] * x[:, :, idx - 3]
    return acc / dt


def store_mask(Nt, store_time_steps):
    if store_time_steps is None:
        return np.ones(Nt, dtype=bool)
    m = np.asarray(store_time_steps)
    if m.dtype == bool:
        return m.copy()
    out = np.zeros(Nt, dtype=bool)
    out[m] = True
    return out


def evolve(circuit, dt=0.05, Nt=1000, f=0.0, Is=0.0, Vs=0.0, T=0.0, store_time_steps=None,
           store_theta=True, store_voltage=True, store_current=True, configs=None,
           stencil_width=3, cp=None, noise=None, solver=superlu_solver):
    """Full reference semantics of ``TimeEvolutionProblem(...).compute()``.

    Returns dict with 'theta', 'current', 'voltage' (each (Nj, W, n_stored) or None)
    and 'store_mask'.
    """
    Nj, Nf = circuit._Nj(), circuit._Nf()
    inputs = Inputs(Nj, Nf, Nt, f, Is, Vs, T)
    W = inputs.W
    mask = store_mask(Nt, store_time_steps)
    if configs is None:
        configs = []
    cfg = list(configs)
    if len(cfg) == 0:
        cfg.append(np.zeros((Nj, W)))
    while len(cfg) < stencil_width - 1:
        cfg.append(cfg[-1])
    cfg = [np.broadcast_to(np.asarray(c, dtype=np.double).reshape(Nj, -1), (Nj, W)) for c in cfg]
    th_mask = mask.copy() if (store_theta or store_voltage) else np.zeros(Nt, dtype=bool)
    I_mask = mask.copy() if (store_current or store_voltage) else np.zeros(Nt, dtype=bool)
    Vth, VI = th_mask.copy(), I_mask.copy()
    t_ids = np.flatnonzero(mask)
    off = stencil_width - 1
    if store_voltage and len(t_ids):
        ids = (t_ids[:, None] - np.arange(off)).ravel()
        ids = ids[(ids >= 0) & (ids < Nt)]
        Vth[ids] = True
        if circuit._has_inductance():
            VI[ids] = True
    th, I = evolve_core(circuit, dt, Nt, inputs, Vth, VI, cfg, stencil_width, cp, noise, solver)
    V = None
    S1 = STENCILS[stencil_width][0]
    if store_voltage:
        ts = np.flatnonzero(mask[Vth])
        V = _derivative(th, ts + off, S1, dt)
        if circuit._has_inductance():
            Vi = _derivative(I, ts + off, S1, dt)
            V = V + (circuit.get_inductance() @ Vi.reshape((Nj, -1))).reshape(V.shape)
        th = np.delete(th, np.flatnonzero((Vth & ~th_mask)[Vth]) + off, axis=2)
        I = np.delete(I, np.flatnonzero((VI & ~I_mask)[VI]) + off, axis=2)
    th = th[:, :, off:]
    I = I[:, :, off:]
    return {"theta": th if store_theta else None, "current": I if store_current else None,
            "voltage": V, "store_mask": mask, "W": W}


def vortex(A, theta):
    """n = -A round(theta / 2pi) per stored step, as integers. theta: (Nj, W, Nt_s)."""
    Nj, W, Ns = theta.shape
    out = np.zeros((A.shape[0], W, Ns), dtype=np.int64)
    for s in range(Ns):
        out[:, :, s] = (-A @ np.round(theta[:, :, s] / (2.0 * np.pi))).astype(int)
    return out


def dense_step(A, Ic, R, C, L, dt, th1, th2, Is, f):
    """One stencil-3 step with a dense solve; brute force for tiny circuits.

    Solves the saddle-point form directly: the new phases theta and cycle
    currents J satisfy  c0 theta + x = A^T J  and  A (theta + L (A^T J)) = -2 pi f,
    i.e. Kirchhoff's current law written in cycle space plus the fluxoid
    constraint, with x = Ic sin(2 th1 - th2) + c1 th1 + c2 th2 - Is.
    """
    A = np.asarray(A.todense() if scipy.sparse.issparse(A) else A, dtype=np.double)
    L = np.asarray(L.todense() if scipy.sparse.issparse(L) else L, dtype=np.double)
    Rv = 1 / (dt * R)
    Cv = C / dt ** 2
    c0, c1, c2 = Rv + Cv, -Rv - 2 * Cv, Cv
    x = (Ic * np.sin(2 * th1 - th2) + c1 * th1 + c2 * th2 - Is)
    Nf, Nj = A.shape
    # unknowns: theta (Nj), J (Nf)
    top = np.hstack((np.diag(c0), -A.T))
    bot = np.hstack((A, A @ L @ A.T))
    M = np.vstack((top, bot))
    rhs = np.concatenate((-x, -2 * np.pi * f))
    sol = np.linalg.solve(M, rhs)
    return sol[:Nj]
