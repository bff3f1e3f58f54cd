"""GPU parity: libjja's HIP path against the reference's own trajectories (golden
fixtures) and against the CPU oracle, through the drop-in API and the C ABI.

Tolerance (FP64): per stored step, max |gpu - ref| <= 1e-9 * max |ref| for theta,
current and voltage; vortex configurations exactly equal.
"""
import numpy as np
import pytest

import pyjjasim_amd as pa
from cases import CASES, T_CASES, build, g, oracle_kwargs, traj
from oracle import te_oracle as to

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def assert_close_per_step(got, ref, rtol=RTOL, what=""):
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    for s in range(ref.shape[2]):
        scale = np.abs(ref[:, :, s]).max()
        err = np.abs(got[:, :, s] - ref[:, :, s]).max()
        assert err <= rtol * max(scale, 1e-300), f"{what} step {s}: err {err:.3e} scale {scale:.3e}"


def _compare_to_golden(case, res):
    for key, getter in (("theta", res.get_theta), ("current", res.get_current),
                        ("voltage", res.get_voltage)):
        name = f"{case}__{key}"
        if name in traj().files:
            assert_close_per_step(getter(), g(case, key), what=f"{case}:{key}")
    np.testing.assert_array_equal(res.get_vortex_configuration(), g(case, "n"))


@pytest.mark.parametrize("case", [c for c in CASES if c not in T_CASES])
def test_gpu_matches_reference_T0(case):
    print(f"[gpu parity] {case}", flush=True)
    c, kw, _ = build(case)
    res = pa.TimeEvolutionProblem(c, **kw).compute()
    _compare_to_golden(case, res)


@pytest.mark.parametrize("case", T_CASES)
def test_gpu_reference_noise_matches_reference(case):
    """T > 0 with the host-replayed legacy RNG stream: the reference's exact draws."""
    print(f"[gpu parity] {case} (reference noise)", flush=True)
    c, kw, seed = build(case)
    np.random.seed(seed)
    res = pa.TimeEvolutionProblem(c, noise="reference", **kw).compute()
    _compare_to_golden(case, res)


def _bench_like(n=12, W=8, Nt=400, T=0.01, seed=99):
    c = pa.SquareArray(n, n)
    rng = np.random.default_rng(4)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    c.set_capacitance(1.0)
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.0, 2.0, W)[None, :, None]
    kw = dict(time_step=0.05, time_step_count=Nt, external_flux=0.25, current_sources=Is,
              temperature=T, store_time_steps=[Nt // 3, Nt - 1], store_current=False,
              store_voltage=False)
    return c, kw


def test_gpu_philox_matches_oracle():
    """Kernel noise (Philox4x32-10 + Box-Muller) equals the oracle's stream."""
    print("[gpu parity] philox bench-like", flush=True)
    c, kw = _bench_like()
    res = pa.TimeEvolutionProblem(c, seed=99, **kw).compute()
    ok = oracle_kwargs(kw)
    ref = to.evolve(c, noise=to.philox_noise(99, 8, c.junction_count()), **ok)
    assert_close_per_step(res.get_theta(), ref["theta"], what="theta")
    np.testing.assert_array_equal(res.get_vortex_configuration(),
                                  to.vortex(c.get_cycle_matrix(), ref["theta"]))


def test_gpu_callable_sources_match_oracle():
    """Callable Is(i) (Shapiro drive) evaluated per chunk on the host."""
    print("[gpu parity] callable source", flush=True)
    c = pa.SquareArray(8, 8)
    base = c.current_base(0.0)
    Idc = np.linspace(0.2, 1.6, 5)
    Is = lambda i: base[:, None] * (Idc + 0.5 * np.sin(0.3 * i * 0.05))
    kw = dict(time_step=0.05, time_step_count=700, external_flux=0.1, current_sources=Is,
              store_time_steps=np.arange(0, 700, 50))
    res = pa.TimeEvolutionProblem(c, steps_per_launch=128, **kw).compute()
    ref = to.evolve(c, **oracle_kwargs(kw))
    for key in ("theta", "current", "voltage"):
        assert_close_per_step(getattr(res, "get_" + key)(), ref[key], what=key)


def test_gpu_deterministic_across_launch_shapes():
    """Same seed -> identical bits for any block size and chunking."""
    print("[gpu parity] determinism", flush=True)
    c, kw = _bench_like(n=10, W=4, Nt=300)
    a = pa.TimeEvolutionProblem(c, seed=5, block_size=1024, **kw).compute().get_theta()
    b = pa.TimeEvolutionProblem(c, seed=5, block_size=128, steps_per_launch=37, **kw).compute().get_theta()
    assert np.array_equal(a, b)


def test_gpu_rejects_custom_cpr():
    c = pa.SquareArray(3, 3)
    cpr = pa.CurrentPhaseRelation(lambda Ic, t: Ic * np.sin(t) ** 3, lambda Ic, t: 0 * t,
                                  lambda Ic, t: 0 * t)
    with pytest.raises(ValueError, match="current-phase relation"):
        pa.TimeEvolutionProblem(c, current_phase_relation=cpr, time_step_count=5).compute()


@pytest.mark.parametrize("T", [0.0, 0.02])
def test_gpu_global_face_store_bit_identical_to_lds(T):
    """The global-memory face store runs the same record program and arithmetic as
    the LDS one: identical bits (the factor program; the lattice solve's two
    stores agree to rounding, tested with the lattice solve)."""
    print(f"[gpu parity] global store T={T}", flush=True)
    c = pa.SquareArray(12, 9)
    rng = np.random.default_rng(8)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    c.set_capacitance(0.7)
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.3, 1.5, 3)[None, :, None]
    out = {}
    for store in ("lds", "global"):
        p = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=400, external_flux=0.2,
                                    current_sources=Is, temperature=T, seed=5, block_size=128,
                                    face_store=store, solver="cholesky")
        r = p.compute()
        out[store] = (r.get_theta(), r.get_voltage(), r.get_vortex_configuration())
    for a, b in zip(out["lds"], out["global"]):
        np.testing.assert_array_equal(a, b)


def test_gpu_large_array_global_store_matches_oracle():
    """SquareArray(160, 160): 25,281 faces, too many for LDS; the global face store
    against the oracle (SuperLU) over a short horizon."""
    print("[gpu parity] 160x160 global store", flush=True)
    c = pa.SquareArray(160, 160)
    rng = np.random.default_rng(2)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    Is = 0.8 * c.current_base(0.0)[:, None, None] * np.array([1.0, 1.4])[None, :, None]
    kw = dict(time_step=0.05, time_step_count=30, external_flux=0.15, current_sources=Is)
    r = pa.TimeEvolutionProblem(c, **kw).compute()
    from oracle import te_oracle
    Nt, Nj, Nf = 30, c.junction_count(), c.face_count()
    inputs = te_oracle.Inputs(Nj, Nf, Nt, 0.15, Is, 0.0, 0.0)
    zero = np.zeros((Nj, 2))
    th_o, _ = te_oracle.evolve_core(c, 0.05, Nt, inputs, np.ones(Nt, dtype=bool), np.zeros(Nt, dtype=bool),
                                    [zero, zero])
    th = r.get_theta()
    scale = np.abs(th_o[:, :, 2:]).max()
    assert np.abs(th - th_o[:, :, 2:]).max() <= 1e-9 * scale
    n_o = te_oracle.vortex(c.get_cycle_matrix(), th_o[:, :, 2:])
    np.testing.assert_array_equal(r.get_vortex_configuration(), n_o)


@pytest.mark.parametrize("parts", [1, 2])
def test_gpu_faceless_circuit_matches_oracle(parts):
    """A circuit without faces (a tree: K is 0 x 0, theta = -x / c0) against the
    oracle, with one and with two workgroups per column."""
    print(f"[gpu parity] faceless circuit, parts={parts}", flush=True)
    g = pa.EmbeddedGraph(np.array([0., 1., 2., 1.]), np.array([0., 0., 0., 1.]),
                         np.array([0, 1, 1]), np.array([1, 2, 3]))
    c = pa.Circuit(g)
    assert c.face_count() == 0
    Is = np.array([0.5, 1.2, 0.3])[:, None, None] * np.array([1.0, 1.5])[None, :, None]
    kw = dict(time_step=0.05, time_step_count=200, current_sources=Is, store_time_steps=[10, 199])
    res = pa.TimeEvolutionProblem(c, block_size=256, parts=parts, **kw).compute()
    ref = to.evolve(c, **oracle_kwargs(kw))
    for key in ("theta", "current", "voltage"):
        assert_close_per_step(getattr(res, "get_" + key)(), ref[key], what=key)


@pytest.mark.parametrize("Nt", [1, 2, 3])
def test_gpu_shortest_horizons_match_oracle(Nt):
    """One, two and three steps, with every step stored (the prefix of initial
    conditions and the first steps of the stencil)."""
    print(f"[gpu parity] Nt={Nt}", flush=True)
    c = pa.SquareArray(5, 4)
    c.set_capacitance(0.7)
    Is = c.current_base(0.0)[:, None, None] * np.array([0.6, 1.4])[None, :, None]
    kw = dict(time_step=0.05, time_step_count=Nt, external_flux=0.2, current_sources=Is)
    res = pa.TimeEvolutionProblem(c, **kw).compute()
    ref = to.evolve(c, **oracle_kwargs(kw))
    for key in ("theta", "current", "voltage"):
        assert_close_per_step(getattr(res, "get_" + key)(), ref[key], what=key)
    np.testing.assert_array_equal(res.get_vortex_configuration(),
                                  to.vortex(c.get_cycle_matrix(), ref["theta"]))


def test_gpu_single_junction_matches_oracle():
    """The smallest circuit: two nodes and one junction (RSJ)."""
    print("[gpu parity] single junction", flush=True)
    g1 = pa.EmbeddedGraph(np.array([0., 1.]), np.array([0., 0.]), np.array([0]), np.array([1]))
    c = pa.Circuit(g1)
    Is = np.array([[[1.3]]])
    kw = dict(time_step=0.05, time_step_count=400, current_sources=Is, store_time_steps=[100, 399])
    res = pa.TimeEvolutionProblem(c, **kw).compute()
    ref = to.evolve(c, **oracle_kwargs(kw))
    for key in ("theta", "current", "voltage"):
        assert_close_per_step(getattr(res, "get_" + key)(), ref[key], what=key)
