"""Static problems on the device (the London approximation and the Newton
iteration, static_problem.py:550-607, 1385-1515) against results the
reference wrote (tests/golden/make_golden.py: static.npz)."""
import os

import numpy as np
import pytest

import pyjjasim_amd as pa

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GETTERS = ("get_phase", "get_current", "get_cycle_current", "get_flux", "get_magnetic_energy",
           "get_josephson_energy", "get_energy", "get_vortex_configuration")


def _cases():
    """The circuits of make_golden._static_cases, built the same way."""
    out = {}
    c = pa.SquareArray(6, 6)
    out["sq6_bias"] = c
    c = pa.SquareArray(7, 7)
    c.set_inductance(0.5)
    out["sq7_L_vortex"] = c
    c = pa.HoneycombArray(5, 5)
    rng = np.random.default_rng(31)
    c.set_critical_current(np.exp(0.2 * rng.standard_normal(c.junction_count())))
    out["hc5_hetero"] = c
    c = pa.TriangularArray(5, 5)
    c.set_inductance(0.4 + 0.3 * rng.random(c.junction_count()))
    out["tri5_selfL"] = c
    out["sq6_overbias"] = pa.SquareArray(6, 6)
    out["sq8_pair"] = pa.SquareArray(8, 8)
    return out


def _rel(got, want):
    return np.abs(np.asarray(got) - want).max() / max(np.abs(want).max(), 1e-300)


@pytest.mark.parametrize("name", ["sq6_bias", "sq7_L_vortex", "hc5_hetero", "tri5_selfL", "sq6_overbias",
                                  "sq8_pair"])
def test_gpu_static_problem_matches_reference(name):
    """The approximation to 1e-9 of its largest value; the Newton status,
    iteration count and is-target flags exactly and its residual history to
    1e-6 wherever the residual is above 1e-9; for converged problems theta and
    every derived quantity to 1e-9 and the Kirchhoff and winding errors below
    the tolerance; stability as the reference decides it."""
    d = np.load(os.path.join(HERE, "golden", "static.npz"), allow_pickle=False)
    g = {k.split("__", 1)[1]: d[k] for k in d.files if k.startswith(name + "__")}
    c = _cases()[name]
    prob = pa.StaticProblem(c, current_sources=g["in_Is"], external_flux=g["in_f"],
                            vortex_configuration=g["in_n"])
    appr = prob.approximate()
    e = _rel(appr.get_theta(), g["approx_theta"])
    print(f"[static] {name}: approximation rel err {e:.2e}", flush=True)
    assert e < 1e-9
    cfg, status, info = prob.compute()
    print(f"  status {status} iterations {info.get_number_of_iterations()} residual {info.get_residual()}",
          flush=True)
    assert status == int(g["status"])
    assert info.get_number_of_iterations() == int(g["iterations"])
    np.testing.assert_array_equal(info.get_is_target_vortex_configuration(), g["is_target_n"])
    big = g["residual"] > 1e-9
    np.testing.assert_allclose(info.get_residual()[big], g["residual"][big], rtol=1e-6)
    assert cfg.is_stable() == int(g["is_stable"])
    if status == 0:
        assert _rel(cfg.get_theta(), g["theta"]) < 1e-9
        for getter in GETTERS:
            err = _rel(getattr(cfg, getter)(), g[getter])
            assert err < 1e-9, (getter, err)
        assert max(cfg.get_error()) < 1e-10
        assert cfg.is_target_solution()
