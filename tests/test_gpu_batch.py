"""Batched problems (TimeEvolutionBatch) and separable time-dependent inputs on
the GPU: a batch returns exactly what each problem's own compute() returns (run
with the batch's solver, and for the lattice solve the batch's layout)."""
import numpy as np
import pytest

import pyjjasim_amd as pa

pytestmark = pytest.mark.gpu


def _problems():
    Nt, dt = 300, 0.05
    t = dt * np.arange(Nt)
    a = pa.SquareArray(9, 7)
    a.set_inductance(1.0)
    pa_ = pa.TimeEvolutionProblem(
        a, time_step=dt, time_step_count=Nt, external_flux=0.3,
        current_sources=pa.SeparableSource(a.current_base(0.0)[:, None] * np.array([0.7, 1.2])[None, :],
                                           1.0 + 0.5 * np.sin(0.3 * t)))
    b = pa.HoneycombArray(5, 4)
    pb = pa.TimeEvolutionProblem(b, time_step=dt, time_step_count=Nt, external_flux=0.1,
                                 current_sources=0.8 * b.current_base(0.0)[:, None, None])
    c = pa.TriangularArray(6, 5)
    rng = np.random.default_rng(4)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    pc = pa.TimeEvolutionProblem(c, time_step=dt, time_step_count=Nt, temperature=0.01, seed=9,
                                 current_sources=c.current_base(0.0)[:, None, None] *
                                 np.linspace(0.5, 1.5, 3)[None, :, None], face_store="global")
    return [pa_, pb, pc]


def _own(p, tb, i):
    """Problem i's own compute() with the face-system solver the batch used for
    it, and for the lattice solve also its layout (the face store and the part
    count set the lattice solve's rounding; the factor's bits do not depend on
    the layout)."""
    saved = (p.solver, p.block_size, p.parts, p.face_store)
    p.solver = tb.solvers[i]
    if tb.solvers[i] == "lattice":
        p.block_size, p.parts, p.face_store = tb.layouts[i]
    try:
        return p.compute()
    finally:
        p.solver, p.block_size, p.parts, p.face_store = saved


def _same(r1, r2):
    np.testing.assert_array_equal(r1.get_theta(), r2.get_theta())
    np.testing.assert_array_equal(r1.get_voltage(), r2.get_voltage())
    np.testing.assert_array_equal(r1.get_vortex_configuration(), r2.get_vortex_configuration())


def test_gpu_batch_matches_individual_runs():
    print("[batch] mixed lattices, LDS and global stores", flush=True)
    probs = _problems()
    tb = pa.TimeEvolutionBatch(probs, block_size=256)
    batch = tb.compute()
    assert tb.solvers == ["lattice", "cholesky", "cholesky"]
    for i, (p, rb) in enumerate(zip(probs, batch)):
        _same(rb, _own(p, tb, i))


def test_gpu_batch_autotune():
    print("[batch] autotune", flush=True)
    probs = _problems()[:2]
    tb = pa.TimeEvolutionBatch(probs, autotune=True, block_sizes=(128, 256), tune_steps=10)
    batch = tb.compute()
    assert (tb.tuning["block_size"], tb.tuning["parts"]) in [(128, 1), (256, 1)] + \
        [(1024, q) for q in (2, 4, 8, 16)]
    assert 100 <= tb.tuning["steps_per_launch"] <= 4000
    for i, (p, rb) in enumerate(zip(probs, batch)):
        _same(rb, _own(p, tb, i))


def test_gpu_separable_source_matches_oracle():
    """A SeparableSource drive equals the materialised (Nj, W, Nt) array in the
    oracle (time_evolution.py:119-130 broadcasting)."""
    print("[batch] separable drive vs oracle", flush=True)
    from oracle import te_oracle
    c = pa.SquareArray(6, 5)
    c.set_capacitance(0.5)
    Nt, dt = 200, 0.05
    drive = pa.SeparableSource(c.current_base(0.0), 0.9 + 0.5 * np.sin(0.3 * dt * np.arange(Nt)))
    r = pa.TimeEvolutionProblem(c, time_step=dt, time_step_count=Nt, external_flux=0.2,
                                current_sources=drive).compute()
    Nj, Nf = c.junction_count(), c.face_count()
    inputs = te_oracle.Inputs(Nj, Nf, Nt, 0.2, np.asarray(drive), 0.0, 0.0)
    zero = np.zeros((Nj, 1))
    th_o, _ = te_oracle.evolve_core(c, dt, Nt, inputs, np.ones(Nt, dtype=bool), np.zeros(Nt, dtype=bool),
                                    [zero, zero])
    th = r.get_theta()
    assert np.abs(th - th_o[:, :, 2:]).max() <= 1e-9 * np.abs(th_o).max()


def test_gpu_batch_split_columns_match_individual_runs():
    """Every column of the batch worked by four workgroups (all face stores
    global): the results of each problem's own compute()."""
    print("[batch] four parts per column", flush=True)
    probs = _problems()
    tb = pa.TimeEvolutionBatch(probs, block_size=256, parts=4)
    batch = tb.compute()
    assert tb.solvers == ["lattice", "cholesky", "cholesky"]
    assert all(lay[1] == 4 and lay[2] == "global" for lay in tb.layouts)
    for i, (p, rb) in enumerate(zip(probs, batch)):
        _same(rb, _own(p, tb, i))


def test_gpu_batch_balanced_parts_match_individual_runs():
    """Problems of one launch with different part counts (the largest split,
    the others whole, all face stores global): each problem's own results."""
    print("[batch] balanced parts per problem", flush=True)
    probs = _problems()
    tb = pa.TimeEvolutionBatch(probs, block_size=256, parts="balanced", max_parts=8)
    batch = tb.compute()
    for i, (p, rb) in enumerate(zip(probs, batch)):
        _same(rb, _own(p, tb, i))


def test_gpu_batch_autotune_short_horizon():
    """Autotuning a batch whose horizon is shorter than the tuning runs keeps
    every trial launch within time_step_count and gives each problem's own
    results; a launch past the horizon is refused."""
    print("[batch] short-horizon autotune", flush=True)
    probs = []
    for n, f in ((6, 0.1), (8, 0.2)):
        c = pa.SquareArray(n, n)
        Is = c.current_base(0.0)[:, None, None] * np.array([0.5, 1.1])[None, :, None]
        probs.append(pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=12, external_flux=f,
                                             current_sources=Is, temperature=0.01, seed=4))
    tb = pa.TimeEvolutionBatch(probs, autotune=True, block_sizes=(64, 128), tune_steps=40)
    for i, (r, p) in enumerate(zip(tb.compute(), probs)):
        _same(r, _own(p, tb, i))
    eng = tb.make_engine()
    try:
        eng.advance(10)
        with pytest.raises(ValueError, match="beyond"):
            eng.launch(3)
    finally:
        eng.close()


def test_gpu_batch_autotune_keeps_reference_noise_stream():
    """Autotuning launches draw from numpy's global RNG; the tuned run starts
    from the caller's RNG state, so with noise="reference" at T > 0 a tuned
    batch of one problem equals that problem's own compute() after the same
    np.random.seed (with several problems the launches draw the problems'
    normals chunk by chunk, an order of their own)."""
    print("[batch] autotune with reference noise", flush=True)

    def problem():
        c = pa.SquareArray(7, 6)
        Is = c.current_base(0.0)[:, None, None] * np.array([0.5, 1.1])[None, :, None]
        return pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=120, external_flux=0.2,
                                       current_sources=Is, temperature=0.02, noise="reference")
    np.random.seed(1234)
    tb = pa.TimeEvolutionBatch([problem()], autotune=True, block_sizes=(64, 128), tune_steps=10)
    batch = tb.compute()
    np.random.seed(1234)
    _same(batch[0], _own(problem(), tb, 0))
