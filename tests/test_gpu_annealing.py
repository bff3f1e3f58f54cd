"""AnnealingProblem on the device (time_evolution.py:1070-1191) against runs of
the reference with numpy's global RNG seeded (tests/golden/annealing.npz): in
noise="reference" mode the device run draws the same normals, so the
temperature profile (each interval's decision from the device-reduced vortex
mobility), the final vortex configurations, the Newton status and the final
phases must be the reference's."""
import os

import numpy as np
import pytest

import pyjjasim_amd as pa

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", ["sq6", "hc4"])
def test_gpu_annealing_matches_reference(name):
    d = np.load(os.path.join(HERE, "golden", "annealing.npz"), allow_pickle=False)
    c, f, count = {"sq6": (pa.SquareArray(6, 6), 0.2, 25), "hc4": (pa.HoneycombArray(4, 4), 0.1, 20)}[name]
    ap = pa.AnnealingProblem(c, time_step=0.5, interval_steps=10, external_flux=f, current_sources=0.0,
                             problem_count=2, interval_count=count, vortex_mobility=0.001, start_T=0.5,
                             T_factor=1.03, noise="reference")
    np.random.seed(2024)
    status, configs, profiles = ap.compute()
    print(f"[annealing] {name}: status {status}, final T {profiles[-1]}", flush=True)
    np.testing.assert_array_equal(profiles, d[name + "__profiles"])
    np.testing.assert_array_equal(np.stack([cfg.get_vortex_configuration() for cfg in configs], axis=1),
                                  d[name + "__n"])
    np.testing.assert_array_equal(status, d[name + "__status"])
    th = np.stack([cfg.get_theta() for cfg in configs], axis=1)
    ok = d[name + "__status"] == 0
    err = np.abs(th[:, ok] - d[name + "__theta"][:, ok]).max() / np.abs(d[name + "__theta"]).max()
    print(f"  converged columns: theta rel err {err:.2e}", flush=True)
    assert err < 1e-9


def test_gpu_annealing_philox_lowers_temperature():
    """With the in-kernel noise stream the same feedback runs: every column's
    temperature follows the reference's rule (each interval multiplies or
    divides it by T_factor), and the final configurations are static
    solutions whose vortex counts sum to the flux quanta the field imposes
    on a small array at f = 0.2 (one vortex per five cells, rounded)."""
    c = pa.SquareArray(6, 6)
    ap = pa.AnnealingProblem(c, time_step=0.5, interval_steps=10, external_flux=0.2, problem_count=3,
                             interval_count=40, vortex_mobility=0.001, start_T=0.5, T_factor=1.03, seed=11)
    status, configs, profiles = ap.compute()
    ratio = profiles[1:] / profiles[:-1]
    assert np.all(np.isclose(ratio, 1.03) | np.isclose(ratio, 1 / 1.03))
    assert np.all(np.isclose(profiles[0], 0.5 * 1.03) | np.isclose(profiles[0], 0.5 / 1.03))
    for cfg, st in zip(configs, status):
        if st == 0:
            assert cfg.is_solution()
            assert abs(cfg.get_vortex_configuration().sum() - round(0.2 * c.face_count())) <= 2
