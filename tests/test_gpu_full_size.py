"""Full-size workloads checked through properties that hold at any size:

- the fluxoid constraint A (theta + L (I - Is)) = -2 pi f, which every step's
  solve enforces (time_evolution.py:560-570);
- current conservation M (I - Is) = 0 (y = A^T J lies in the cycle space);
- device vortex counts equal to -A rint(theta / 2 pi) of the returned phases
  (time_evolution.py:734-755).

Workloads: the bench's SquareArray(100,100) with W=256 bias points (LDS face
store), and the large SquareArray(512,512) with W=16, beta_L=0.5 and disordered
flux (global face store, split columns)."""
import numpy as np
import pytest

import pyjjasim_amd as pa

pytestmark = pytest.mark.gpu


def _check(c, res, Is, f):
    A, M, L = c.get_cycle_matrix(), c.get_cut_matrix(), c.get_inductance()
    th, I = res.get_theta(), res.get_current()
    n = res.get_vortex_configuration()
    f = np.broadcast_to(np.asarray(f, dtype=np.float64).reshape(-1, 1), (c.face_count(), 1))
    for s in range(th.shape[2]):
        d = I[:, :, s] - Is
        lhs = A @ (th[:, :, s] + L @ d) + 2 * np.pi * f
        scale = 1.0 + np.abs(th[:, :, s]).max()
        assert np.abs(lhs).max() <= 1e-9 * scale, f"fluxoid residual {np.abs(lhs).max():.3e} at slot {s}"
        assert np.abs(M @ d).max() <= 1e-9 * (1.0 + np.abs(d).max()), "current conservation"
        want = (-A @ np.round(th[:, :, s] / (2 * np.pi))).astype(np.int64)
        np.testing.assert_array_equal(n[:, :, s], want)


def test_gpu_bench_workload_invariants():
    """SquareArray(100,100), W=256 bias points 0..2, log-normal Ic, beta_C=1,
    f=0.25, T=0.01 (the bench workload), 120 steps."""
    print("[full size] SquareArray(100,100) W=256", flush=True)
    c = pa.SquareArray(100, 100)
    rng = np.random.default_rng(20261202)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    c.set_capacitance(1.0)
    W, Nt = 256, 120
    Is = c.current_base(0.0)[:, None] * np.linspace(0.0, 2.0, W)[None, :]
    res = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=Nt, external_flux=0.25,
                                  current_sources=Is[:, :, None], temperature=0.01, seed=12345,
                                  store_time_steps=[Nt // 2, Nt - 1]).compute()
    _check(c, res, Is, 0.25)


def test_gpu_large_array_invariants():
    """SquareArray(512,512), W=16 bias points 0.2..1.2, beta_C=1, beta_L=0.5,
    f = 0.1 + 0.02 z per face, T=0.05, 30 steps (global face store, split
    columns)."""
    print("[full size] SquareArray(512,512) W=16", flush=True)
    c = pa.SquareArray(512, 512)
    c.set_capacitance(1.0)
    c.set_inductance(0.5)
    W, Nt = 16, 30
    f = 0.1 + 0.02 * np.random.default_rng(505).standard_normal(c.face_count())
    Is = c.current_base(0.0)[:, None] * np.linspace(0.2, 1.2, W)[None, :]
    prob = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=Nt, external_flux=f[:, None, None],
                                   current_sources=Is[:, :, None], temperature=0.05, seed=7,
                                   store_time_steps=[Nt - 1])
    res = prob.compute()
    _check(c, res, Is, f)
