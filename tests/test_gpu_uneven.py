"""The uneven batch (many problems of different lattices and sizes in one launch
per chunk) against the oracle, and the full 64-problem batch through
size-independent invariants.

Workload family: square, honeycomb and triangular lattices, junction
self-inductance beta_L = 1, R = 1, C = 0, f uniform in [0, 0.5] per problem,
Shapiro drive Is(t) = (Idc + 0.5 sin(0.3 t)) current_base(0) with Idc uniform
in [0.5, 1.5], T = 0, dt = 0.05, W = 1 per problem (time_evolution.py:461-582
for each problem)."""
import os
import sys
import time

import numpy as np
import pytest

import pyjjasim_amd as pa

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _reduced_batch(Nt, seed=77):
    """Eight seeded problems: every lattice, counts 16..40, the Shapiro drive as a
    SeparableSource; theta and current stored at three steps."""
    rng = np.random.default_rng(seed)
    kinds = [pa.SquareArray, pa.HoneycombArray, pa.TriangularArray]
    dt = 0.05
    wave = np.sin(0.3 * dt * np.arange(Nt))
    store = [Nt // 3, 2 * Nt // 3, Nt - 1]
    probs = []
    for i in range(8):
        K = kinds[i % 3]
        n = int(rng.integers(16, 41))
        c = K(n, n)
        c.set_inductance(1.0)
        f = float(rng.uniform(0.0, 0.5))
        Idc = float(rng.uniform(0.5, 1.5))
        drive = pa.SeparableSource(c.current_base(0.0), Idc + 0.5 * wave)
        probs.append((pa.TimeEvolutionProblem(c, time_step=dt, time_step_count=Nt, external_flux=f,
                                              current_sources=drive, store_time_steps=store,
                                              store_voltage=False), f, drive))
    return probs


def _oracle_check(prob, f, drive, res):
    from oracle import te_oracle as to
    c = prob.circuit
    ref = to.evolve(c, dt=prob._dt(), Nt=prob._Nt(), f=f, Is=np.asarray(drive),
                    store_time_steps=prob.store_time_steps, store_voltage=False)
    th, I = res.get_theta(), res.get_current()
    err_th = np.abs(th - ref["theta"]).max() / np.abs(ref["theta"]).max()
    err_I = np.abs(I - ref["current"]).max() / np.abs(ref["current"]).max()
    assert err_th <= 1e-9, f"theta rel err {err_th:.3e}"
    assert err_I <= 1e-9, f"current rel err {err_I:.3e}"
    np.testing.assert_array_equal(res.get_vortex_configuration(),
                                  to.vortex(c.get_cycle_matrix(), ref["theta"]))
    return err_th


@pytest.mark.parametrize("layout", ["balanced", "mixed"])
def test_gpu_uneven_batch_matches_oracle(layout):
    """Eight problems in one plan, each against the oracle on the same inputs:
    theta and current within 1e-9 of the largest value, vortex counts exact.
    "balanced": a part count per problem (all face stores global); "mixed": one
    workgroup per column, four problems with the LDS face store and four with
    the global one in the same launch."""
    Nt = 300
    items = _reduced_batch(Nt)
    probs = [p for p, _, _ in items]
    if layout == "mixed":
        for i, p in enumerate(probs):
            p.face_store = "global" if i % 2 else "lds"
        tb = pa.TimeEvolutionBatch(probs, block_size=256)
    else:
        tb = pa.TimeEvolutionBatch(probs, block_size=256, parts="balanced", max_parts=8)
    print(f"[uneven] {layout}: {[p.circuit.junction_count() for p in probs]} junctions", flush=True)
    results = tb.compute()
    for (p, f, drive), r in zip(items, results):
        err = _oracle_check(p, f, drive, r)
        print(f"  {type(p.circuit).__name__} Nj={p.circuit.junction_count()}: theta rel err {err:.2e}",
              flush=True)


@pytest.mark.timeout(900)
def test_gpu_uneven_batch_full_size_invariants():
    """The full 64-problem batch (seeded lattices with 16..256 cells per side,
    6.0e6 junctions) for 200 steps with balanced parts: per problem the fluxoid
    constraint A (theta + L (I - Is)) = -2 pi f, current conservation
    M (I - Is) = 0 and device vortex counts equal to -A rint(theta / 2 pi)."""
    sys.path.insert(0, ROOT)
    import bench
    Nt = 200
    t0 = time.perf_counter()
    probs, names = bench.uneven_problems(pa, Nt)
    for p in probs:
        p.store_current = True
    print(f"[uneven] 64 problems built in {time.perf_counter() - t0:.0f} s", flush=True)
    t0 = time.perf_counter()
    results = pa.TimeEvolutionBatch(probs, parts="balanced").compute()
    print(f"[uneven] plans and {Nt} steps in {time.perf_counter() - t0:.0f} s", flush=True)
    for p, r, name in zip(probs, results, names):
        c = p.circuit
        A, M, L = c.get_cycle_matrix(), c.get_cut_matrix(), c.get_inductance()
        th, I = r.get_theta()[:, :, -1], r.get_current()[:, :, -1]
        Is = p._Is(Nt - 1)
        d = I - Is
        lhs = A @ (th + L @ d) + 2 * np.pi * p._f(Nt - 1)
        assert np.abs(lhs).max() <= 1e-9 * (1.0 + np.abs(th).max()), name
        assert np.abs(M @ d).max() <= 1e-9 * (1.0 + np.abs(d).max()), name
        want = (-A @ np.round(th / (2 * np.pi))).astype(np.int64)
        np.testing.assert_array_equal(r.get_vortex_configuration()[:, :, -1], want)
