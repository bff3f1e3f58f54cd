"""Pins of the CPU oracle (oracle/te_oracle.py, oracle/philox.py).

* the reference's own trajectories (tests/golden, written by running the
  reference): the oracle must reproduce them bit for bit, since it uses the same
  library calls in the same order;
* Philox4x32-10 against the published Random123 known-answer vectors;
* closed forms and invariants that a dropped term, a wrong sign or a transposed
  operand would break: the fluxoid constraint, current conservation, the
  telescoping time-averaged voltage, the zero-input fixed point, a dense
  brute-force solve of one step, and the noise variance.
"""
import numpy as np
import pytest

import pyjjasim_amd as pa
from cases import CASES, build, g, oracle_kwargs
from oracle import te_oracle as to
from oracle.philox import normals, philox4x32_10


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_reference(case):
    c, kw, seed = build(case)
    if seed is not None:
        np.random.seed(seed)
    r = to.evolve(c, **oracle_kwargs(kw))
    for key in ("theta", "current", "voltage"):
        name = f"{case}__{key}"
        from cases import traj
        if name in traj().files:
            np.testing.assert_array_equal(r[key], g(case, key))
    if r["theta"] is not None:
        np.testing.assert_array_equal(to.vortex(c.get_cycle_matrix(), r["theta"]), g(case, "n"))


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32_10."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        got = philox4x32_10(*ctr, *key)
        assert tuple(int(v) for v in got) == want


def test_philox_normals_statistics():
    z = normals(12345, 7, np.arange(64), 4096)
    assert z.shape == (4096, 64)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    # different steps / columns / seeds give independent draws
    assert abs(np.corrcoef(z.ravel(), normals(12345, 8, np.arange(64), 4096).ravel())[0, 1]) < 0.01
    assert not np.array_equal(normals(1, 7, [0], 16), normals(2, 7, [0], 16))


def _run(c, **kw):
    return to.evolve(c, **kw)


def test_fluxoid_and_current_conservation():
    """A (theta + L (I - Is)) = -2 pi f at every step; M (I - Is) = 0
    (y = A^T J lies in the cycle space)."""
    c = pa.SquareArray(6, 5)
    c.set_inductance(0.6)
    c.set_capacitance(0.4)
    Is = c.current_base(0.0)[:, None, None] * np.array([0.4, 1.3])[None, :, None]
    f = 0.17
    r = _run(c, dt=0.05, Nt=300, f=f, Is=Is)
    A, M, L = c.get_cycle_matrix(), c.get_cut_matrix(), c.get_inductance()
    th, I = r["theta"], r["current"]
    Isb = np.broadcast_to(Is[:, :, 0][:, :, None], I.shape)
    for s in range(0, 300, 37):
        lhs = A @ (th[:, :, s] + L @ (I[:, :, s] - Isb[:, :, s]))
        np.testing.assert_allclose(lhs, -2 * np.pi * f, atol=1e-9)
        np.testing.assert_allclose(M @ (I[:, :, s] - Isb[:, :, s]), 0, atol=1e-10)


def test_time_averaged_voltage_telescopes():
    c = pa.SquareArray(5, 5)
    Is = 1.3 * c.current_base(0.0)[:, None, None]
    Nt, dt = 400, 0.05
    r = _run(c, dt=dt, Nt=Nt, f=0.1, Is=Is)
    Vmean = r["voltage"].mean(axis=2)
    np.testing.assert_allclose(Vmean, (r["theta"][:, :, -1] - 0.0) / (Nt * dt), rtol=1e-9, atol=1e-12)


def test_zero_input_fixed_point():
    c = pa.TriangularArray(4, 4)
    r = _run(c, dt=0.05, Nt=50)
    assert np.all(r["theta"] == 0) and np.all(r["current"] == 0)


def test_single_step_matches_dense_saddle_point_solve():
    rng = np.random.default_rng(3)
    c = pa.HoneycombArray(3, 2)
    Nj = c.junction_count()
    Ic = 0.5 + rng.random(Nj)
    R = 0.5 + rng.random(Nj)
    C = rng.random(Nj)
    c.set_critical_current(Ic)
    c.set_resistance(R)
    c.set_capacitance(C)
    c.set_inductance(0.3 + rng.random(Nj))
    th1, th2 = rng.standard_normal(Nj), rng.standard_normal(Nj)
    Is = rng.standard_normal(Nj) * 0.3
    f = rng.random(c.face_count())
    r = _run(c, dt=0.07, Nt=1, f=f[:, None, None], Is=Is[:, None, None],
             configs=[th1[:, None], th2[:, None]], store_voltage=False)
    want = to.dense_step(c.get_cycle_matrix(), Ic, R, C, c.get_inductance(), 0.07, th1, th2, Is, f)
    np.testing.assert_allclose(r["theta"][:, 0, 0], want, rtol=1e-11, atol=1e-11)


def test_rsj_single_junction_limit():
    """f=0 uniform bias on a square array: every horizontal junction is an RSJ with
    <V> -> sqrt(I^2 - Ic^2) for I > Ic as dt -> 0 (loose check, dt=0.01)."""
    c = pa.SquareArray(3, 2)
    I = 1.5
    Is = I * c.current_base(0.0)[:, None, None]
    Nt, dt = 6000, 0.01
    r = _run(c, dt=dt, Nt=Nt, Is=Is, store_time_steps=[2000, Nt - 1], store_voltage=False,
             store_current=False)
    hor = np.isclose(c.current_base(0.0), 1.0)
    V = (r["theta"][hor, 0, 1] - r["theta"][hor, 0, 0]) / (dt * (Nt - 1 - 2000))
    np.testing.assert_allclose(V, np.sqrt(I ** 2 - 1), rtol=0.02)


def test_noise_variance_matches_closed_form():
    """eta = sqrt(2 T Rv) Z (paper App. D eq. 20; time_evolution.py:538): with Ic = 0,
    C = 0, f = 0 the phase increment per step of an isolated loop is linear in eta."""
    c = pa.SquareArray(2, 2)
    c.set_critical_current(0.0)
    T, dt = 0.3, 0.05
    z = {}

    def noise(i):
        z[i] = np.random.default_rng(i).standard_normal((c.junction_count(), 1))
        return z[i]
    r = _run(c, dt=dt, Nt=3, T=T, noise=noise, store_voltage=False, store_current=False)
    # theta_n = P (theta_{n-1} ... ) with Ic = 0, R = 1: x = eta - Rv th1, th_n = th1 - D P' eta
    Rv = 1 / dt
    th0 = r["theta"][:, 0, 0]
    eta = np.sqrt(2 * T * Rv) * z[0][:, 0]
    A = np.asarray(c.get_cycle_matrix().todense(), dtype=float)
    # y = A^T (A A^T / c0)^-1 A (x / c0); theta = (y - x)/c0 with c0 = Rv
    x = eta
    y = A.T @ np.linalg.solve(A @ A.T / Rv, A @ (x / Rv))
    np.testing.assert_allclose(th0, (y - x) / Rv, rtol=1e-12)


@pytest.mark.parametrize("case", ["custom", "custom_dangling", "hole_8", "squid"])
def test_oracle_reproduces_reference_on_general_graphs(case):
    """The reference's runs on graphs that are not lattices
    (graph_trajectories.npz): the oracle, on this package's graph and circuit,
    reproduces theta, current, voltage and vortex counts bit for bit."""
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    d = np.load(os.path.join(here, "golden", "graph_trajectories.npz"), allow_pickle=False)
    x9, y9 = [0, 1, 2, 0, 1, 2, 0, 1, 2], [0, 0, 0, 1, 1, 1, 2, 2, 2]
    n1, n2 = [4, 2, 7, 0, 5, 6, 4, 1, 8, 7, 7], [3, 1, 6, 3, 2, 3, 5, 0, 5, 4, 8]
    if case == "custom":
        c = pa.Circuit(pa.EmbeddedGraph(x9, y9, n1, n2))
    elif case == "custom_dangling":
        c = pa.Circuit(pa.EmbeddedGraph(x9 + [0.4], y9 + [1.6], n1 + [9], n2 + [3]))
    elif case == "hole_8":
        full = pa.SquareArray(8, 8)
        fx, fy = full.get_node_coordinates()
        c = full.remove_nodes((fx > 2) & (fx < 5) & (fy > 2) & (fy < 5))
    else:
        c = pa.SQUID()
    c.set_capacitance(0.5)
    r = to.evolve(c, dt=float(d[case + "__in_dt"]), Nt=400, f=0.2, Is=d[case + "__in_Is"],
                  store_time_steps=d[case + "__in_ts"])
    for key in ("theta", "current", "voltage"):
        np.testing.assert_array_equal(r[key], d[f"{case}__{key}"])
    np.testing.assert_array_equal(to.vortex(c.get_cycle_matrix(), r["theta"]), d[case + "__n"])


@pytest.mark.parametrize("case", ["harm2", "harm3"])
def test_oracle_reproduces_reference_with_sine_series_cpr(case):
    """The reference's runs with a custom sine-series current-phase relation
    (custom_cpr.npz): the oracle with FourierCPR's evaluation reproduces
    theta, current and voltage bit for bit."""
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    d = np.load(os.path.join(here, "golden", "custom_cpr.npz"), allow_pickle=False)
    cpr = pa.FourierCPR(d[case + "__coef"])
    c = pa.SquareArray(7, 6)
    c.set_capacitance(0.5)
    r = to.evolve(c, dt=0.05, Nt=400, f=0.15, Is=d[case + "__in_Is"], store_time_steps=d[case + "__in_ts"],
                  cp=cpr.eval)
    for key in ("theta", "current", "voltage"):
        np.testing.assert_array_equal(r[key], d[f"{case}__{key}"])


def test_fourier_cpr_is_consistent():
    """FourierCPR([1]) is the default relation; d_eval is the derivative and
    i_eval the integral (from 0) of eval, checked by central differences and
    the trapezoid rule."""
    th = np.linspace(-4, 4, 2001)
    Ic = np.array([1.3])
    one = pa.FourierCPR([1.0])
    dflt = pa.DefaultCPR()
    np.testing.assert_array_equal(one.eval(Ic, th), dflt.eval(Ic, th))
    np.testing.assert_array_equal(one.i_eval(Ic, th), dflt.i_eval(Ic, th))
    cpr = pa.FourierCPR([0.8, -0.25, 0.1])
    h = 1e-6
    fd = (cpr.eval(Ic, th + h) - cpr.eval(Ic, th - h)) / (2 * h)
    assert np.abs(fd - cpr.d_eval(Ic, th)).max() < 1e-8
    x = np.linspace(0, 2.5, 200001)
    y = cpr.eval(Ic, x)
    integral = np.sum((y[1:] + y[:-1]) * np.diff(x)) / 2
    assert abs(integral - cpr.i_eval(Ic, np.array([2.5]))[0]) < 1e-9
