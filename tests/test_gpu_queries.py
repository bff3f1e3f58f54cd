"""Derived result quantities on the device (phase, potential, energies,
supercurrent, cycle current, flux; time_evolution.py:692-983) and trajectories
on graphs that are not lattices, against outputs the reference itself wrote
(tests/golden/make_golden.py: queries.npz, graph_trajectories.npz)."""
import os

import numpy as np
import pytest

import pyjjasim_amd as pa

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GETTERS = ("get_phase", "get_potential", "get_josephson_energy", "get_magnetic_energy",
           "get_capacitive_energy", "get_energy", "get_supercurrent", "get_cycle_current", "get_flux")


def _load(name):
    data = np.load(os.path.join(HERE, "golden", name), allow_pickle=False)
    cases = {}
    for key in data.files:
        case, field = key.split("__", 1)
        cases.setdefault(case, {})[field] = data[key]
    return cases


def _close(got, want, what, rel=1e-9):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want).max() / scale
    assert got.shape == want.shape, (what, got.shape, want.shape)
    assert err <= rel, f"{what}: rel err {err:.3e}"
    return err


def _query_problem(name, g):
    if name == "sq8_L":
        c = pa.SquareArray(8, 8)
        c.set_inductance(1.0)
        c.set_capacitance(1.0)
        return pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=300, external_flux=0.1,
                                       current_sources=g["in_Is"], store_time_steps=g["in_ts"])
    c = pa.SquareArray(7, 6)
    c.set_critical_current(g["in_Ic"])
    c.set_resistance(g["in_R"])
    c.set_capacitance(g["in_C"])
    return pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=200, external_flux=g["in_f"],
                                   current_sources=g["in_Is"], store_time_steps=g["in_ts"])


@pytest.mark.parametrize("case", ["sq8_L", "sq76_hetero"])
def test_gpu_result_queries_match_reference(case):
    """Every derived getter, over all stored steps and over a selection of them,
    equals the reference's to 1e-9 of its largest value (the phases themselves
    agree to the rounding of the two solvers)."""
    g = _load("queries.npz")[case]
    res = _query_problem(case, g).compute()
    print(f"[queries] {case}", flush=True)
    _close(res.get_theta(), g["theta"], "theta")
    for name in GETTERS:
        e = _close(getattr(res, name)(), g[name], name)
        _close(getattr(res, name)(g["select"]), g[name + "__sel"], name + " (selection)")
        print(f"  {name}: rel err {e:.2e}", flush=True)


def test_gpu_result_query_errors():
    """The reference's exceptions: quantities of an output that was not stored
    raise its error, except the energies of an element the circuit lacks (zeros);
    a step that was not stored raises DataAtTimepointNotStored."""
    c = pa.SquareArray(4, 4)
    prob = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=20, external_flux=0.1,
                                   current_sources=0.5 * c.current_base(0.0)[:, None, None],
                                   store_time_steps=[5, 19], store_current=False, store_voltage=False)
    res = prob.compute()
    with pytest.raises(pa.CurrentNotStored):
        res.get_cycle_current()
    with pytest.raises(pa.CurrentNotStored):
        res.get_flux()
    with pytest.raises(pa.VoltageNotStored):
        res.get_potential()
    assert not res.get_magnetic_energy().any()       # no inductance: zeros
    assert not res.get_capacitive_energy().any()     # no capacitance: zeros
    mask = np.zeros(20, dtype=bool)
    mask[6] = True
    with pytest.raises(pa.DataAtTimepointNotStored):
        res.get_phase(mask)


@pytest.mark.parametrize("case", ["custom", "custom_dangling", "hole_8", "squid"])
def test_gpu_general_graph_trajectories_match_reference(case):
    """Time evolution on graphs that are not lattices (a custom graph with
    reversed edges, the same with a dangling junction, a lattice with a hole, the
    SQUID, at a time step short enough for its Ic = 1000 junctions), with a
    current injected at one node and taken out at another:
    theta, current and voltage within 1e-9 of the reference's largest value and
    vortex counts exact, at every 10th of 400 steps."""
    g = _load("graph_trajectories.npz")[case]
    x = [0, 1, 2, 0, 1, 2, 0, 1, 2]
    y = [0, 0, 0, 1, 1, 1, 2, 2, 2]
    n1 = [4, 2, 7, 0, 5, 6, 4, 1, 8, 7, 7]
    n2 = [3, 1, 6, 3, 2, 3, 5, 0, 5, 4, 8]
    if case == "custom":
        c = pa.Circuit(pa.EmbeddedGraph(x, y, n1, n2))
    elif case == "custom_dangling":
        c = pa.Circuit(pa.EmbeddedGraph(x + [0.4], y + [1.6], n1 + [9], n2 + [3]))
    elif case == "hole_8":
        full = pa.SquareArray(8, 8)
        fx, fy = full.get_node_coordinates()
        c = full.remove_nodes((fx > 2) & (fx < 5) & (fy > 2) & (fy < 5))
    else:
        c = pa.SQUID()
    xn, yn = c.get_node_coordinates()
    node = np.zeros(c.node_count())
    node[np.argmin(xn + yn)] = 1.0
    node[np.argmax(xn + yn)] = -1.0
    base = pa.node_to_junction_current(c, node)
    np.testing.assert_allclose(base, g["in_base"], rtol=0, atol=1e-13)
    c.set_capacitance(0.5)
    res = pa.TimeEvolutionProblem(c, time_step=float(g["in_dt"]), time_step_count=400, external_flux=0.2,
                                  current_sources=g["in_Is"], store_time_steps=g["in_ts"]).compute()
    print(f"[graphs] {case}: Nj={c.junction_count()} Nf={c.face_count()}", flush=True)
    _close(res.get_theta(), g["theta"], "theta")
    _close(res.get_current(), g["current"], "current")
    _close(res.get_voltage(), g["voltage"], "voltage")
    np.testing.assert_array_equal(res.get_vortex_configuration(), g["n"])


def test_gpu_streamed_outputs_equal_resident_outputs():
    """Stored steps beyond the device budget stream to the host through rings of
    output slots: a run with a budget of a few steps' worth returns exactly the
    theta, current, voltage and vortex counts of a run with everything resident."""
    print("[streaming] ring of output slots", flush=True)
    c = pa.SquareArray(9, 8)
    c.set_inductance(0.5)
    W, Nt = 3, 700
    Is = c.current_base(0.0)[:, None, None] * np.array([0.4, 0.9, 1.5])[None, :, None]

    def run(budget):
        return pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=Nt, external_flux=0.15,
                                       current_sources=Is, temperature=0.01, seed=3,
                                       store_time_steps=np.arange(0, Nt, 3), steps_per_launch=97,
                                       max_resident_output_bytes=budget).compute()
    row = 8 * W * c.junction_count()
    streamed, resident = run(7 * row), run(None)
    for g in ("get_theta", "get_current", "get_voltage", "get_vortex_configuration"):
        np.testing.assert_array_equal(getattr(streamed, g)(), getattr(resident, g)())


@pytest.mark.parametrize("case", ["harm2", "harm3"])
def test_gpu_sine_series_cpr_matches_reference(case):
    """A sine-series current-phase relation on the device (FourierCPR) against the
    reference run with the same relation as a custom CurrentPhaseRelation
    (custom_cpr.npz): theta, current, voltage, Josephson energy and supercurrent
    within 1e-9 of the reference's largest value, vortex counts exact."""
    g = _load("custom_cpr.npz")[case]
    c = pa.SquareArray(7, 6)
    c.set_capacitance(0.5)
    res = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=400, external_flux=0.15,
                                  current_sources=g["in_Is"], store_time_steps=g["in_ts"],
                                  current_phase_relation=pa.FourierCPR(g["coef"])).compute()
    print(f"[cpr] {case}: coefficients {g['coef']}", flush=True)
    for name, got in (("theta", res.get_theta()), ("current", res.get_current()), ("voltage", res.get_voltage()),
                      ("get_josephson_energy", res.get_josephson_energy()),
                      ("get_supercurrent", res.get_supercurrent())):
        print(f"  {name}: rel err {_close(got, g[name], name):.2e}", flush=True)
    np.testing.assert_array_equal(res.get_vortex_configuration(), g["n"])
