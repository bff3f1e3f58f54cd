"""The lattice solve on the device (the face system of a uniform rectangular
lattice by sine transforms and tridiagonal solves, in place of the staged
factor): trajectories against the oracle (SuperLU, time_evolution.py:505-506,
560-569) at 1e-9 with exact vortex counts, against the factor path on the
device, bit for bit across block sizes and face stores, and gradients of the
adjoint against the factor path's."""
import numpy as np
import pytest

import pyjjasim_amd as pa
from oracle import te_oracle as to
from pyjjasim_amd.solver_plan import CircuitPlan

pytestmark = pytest.mark.gpu


def _lattices():
    out = {}
    c = pa.SquareArray(10, 10)
    out["sq10"] = c
    c = pa.SquareArray(13, 8)
    c.set_inductance(0.6)
    c.set_capacitance(0.4)
    out["selfL_13x8"] = c
    c = pa.SquareArray(6, 17)
    x1, y1, x2, y2 = c.get_junction_coordinates()
    c.set_capacitance(np.where(y1 == y2, 1.0, 2.0))
    rng = np.random.default_rng(8)
    c.set_critical_current(np.exp(0.2 * rng.standard_normal(c.junction_count())))
    out["aniso_6x17"] = c
    return out


def _inputs(c, W=3):
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.3, 1.6, W)[None, :, None]
    return dict(time_step=0.05, external_flux=0.21, current_sources=Is)


@pytest.mark.parametrize("name", sorted(_lattices()))
def test_gpu_lattice_solve_matches_oracle(name):
    c = _lattices()[name]
    Nt, kw = 400, _inputs(c)
    ts = [0, 199, Nt - 1]
    print(f"[lattice] {name} vs oracle", flush=True)
    prob = pa.TimeEvolutionProblem(c, time_step_count=Nt, store_time_steps=ts, solver="lattice", **kw)
    res = prob.compute()
    ref = to.evolve(c, dt=0.05, Nt=Nt, f=0.21, Is=kw["current_sources"], store_time_steps=ts)
    th, V = res.get_theta(), res.get_voltage()
    assert np.abs(th - ref["theta"]).max() <= 1e-9 * np.abs(ref["theta"]).max()
    assert np.abs(V - ref["voltage"]).max() <= 1e-9 * np.abs(ref["voltage"]).max()
    np.testing.assert_array_equal(res.get_vortex_configuration(), to.vortex(c.get_cycle_matrix(), ref["theta"]))


def test_gpu_lattice_solve_matches_factor_path_at_T():
    """Philox noise at T > 0, 2000 steps: lattice and factor paths agree to
    1e-9 of the phase scale with identical vortex counts."""
    c = _lattices()["selfL_13x8"]
    Nt, kw = 2000, _inputs(c, W=8)
    ts = [999, Nt - 1]
    out = {}
    for solver in ("lattice", "cholesky"):
        print(f"[lattice] T > 0, {solver}", flush=True)
        out[solver] = pa.TimeEvolutionProblem(c, time_step_count=Nt, store_time_steps=ts, temperature=0.03,
                                              seed=17, solver=solver, **kw).compute()
    a, b = out["lattice"].get_theta(), out["cholesky"].get_theta()
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
    np.testing.assert_array_equal(out["lattice"].get_vortex_configuration(),
                                  out["cholesky"].get_vortex_configuration())


def test_gpu_lattice_solve_bit_identical_across_layouts():
    """Each output of the transforms is summed by one thread (LDS store) or one
    matrix-core chain (global store) in a fixed order and each mode's
    recurrence runs in one thread: block size and launch length do not change
    a bit. The two stores round their transforms differently and agree to
    1e-12."""
    c = _lattices()["aniso_6x17"]
    Nt, kw = 150, _inputs(c)
    runs = {}
    for bs, store, spl in ((64, "lds", None), (256, "lds", 37), (1024, "lds", None),
                           (256, "global", 50), (64, "global", None), (1024, "global", 29)):
        print(f"[lattice] block {bs}, {store} store", flush=True)
        r = pa.TimeEvolutionProblem(c, time_step_count=Nt, store_time_steps=[Nt // 2, Nt - 1], temperature=0.01,
                                    seed=4, block_size=bs, face_store=store, steps_per_launch=spl,
                                    solver="lattice", **kw).compute()
        runs.setdefault(store, []).append(r.get_theta())
    for store in runs:
        for th in runs[store][1:]:
            np.testing.assert_array_equal(th, runs[store][0])
    a, b = runs["global"][0], runs["lds"][0]
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_gpu_lattice_adjoint_matches_factor_path():
    """Gradients in Ic, f and Is through the lattice solve equal the factor
    path's to 1e-9 (the adjoint solves the same symmetric K)."""
    torch = pytest.importorskip("torch")
    c = _lattices()["selfL_13x8"]
    Nj, Nf, W, Nt = c.junction_count(), c.face_count(), 3, 120
    rng = np.random.default_rng(2)
    Ic = torch.tensor(np.exp(0.1 * rng.standard_normal(Nj)), dtype=torch.float64, device="cuda")
    f = torch.tensor(0.2 + 0.05 * rng.standard_normal((Nf, W)), dtype=torch.float64, device="cuda")
    Is = torch.tensor(np.outer(c.current_base(0.0), [0.4, 0.9, 1.3]), dtype=torch.float64, device="cuda")
    store = np.zeros(Nt, dtype=bool)
    store[[Nt // 2, Nt - 1]] = True
    G = torch.tensor(rng.standard_normal((Nj, W, 2)), dtype=torch.float64, device="cuda")
    grads = {}
    for solver in ("lattice", "cholesky"):
        print(f"[lattice] adjoint, {solver}", flush=True)
        with pa.DifferentiableEvolution(c, time_step=0.05, time_step_count=Nt, problem_count=W, temperature=0.01,
                                        store_time_steps=store, checkpoint_every=40, seed=6,
                                        solver=solver) as evo:
            assert evo.eng.host.uses_lattice == (solver == "lattice")
            leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
            (evo(*leaves) * G).sum().backward()
            grads[solver] = [t.grad.cpu().numpy() for t in leaves]
    for a, b in zip(grads["lattice"], grads["cholesky"]):
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def test_gpu_split_lattice_matches_oracle_and_unsplit():
    """A column split over 2 and 4 workgroups runs the lattice solve between
    column barriers; each mode's recurrence takes more blocks (B grows with the
    parts), which changes only the rounding: trajectories agree with the
    unsplit lattice solve to 1e-12 and with the oracle to 1e-9, with exact
    vortex counts."""
    c = pa.SquareArray(60, 40)
    Nt, kw = 300, _inputs(c, W=2)
    ts = [149, Nt - 1]
    ref = to.evolve(c, dt=0.05, Nt=Nt, f=0.21, Is=kw["current_sources"], store_time_steps=ts)
    runs = {}
    for parts in (1, 2, 4):
        print(f"[lattice] split SquareArray(60,40), parts {parts}", flush=True)
        prob = pa.TimeEvolutionProblem(c, time_step_count=Nt, store_time_steps=ts, solver="lattice",
                                       block_size=512, face_store="global", parts=parts, **kw)
        assert CircuitPlan(c, 0.05, block_size=512, face_store="global", parts=parts).uses_lattice
        res = prob.compute()
        runs[parts] = res
        th, V = res.get_theta(), res.get_voltage()
        assert np.abs(th - ref["theta"]).max() <= 1e-9 * np.abs(ref["theta"]).max()
        assert np.abs(V - ref["voltage"]).max() <= 1e-9 * np.abs(ref["voltage"]).max()
        np.testing.assert_array_equal(res.get_vortex_configuration(), to.vortex(c.get_cycle_matrix(), ref["theta"]))
    for parts in (2, 4):
        a, b = runs[parts].get_theta(), runs[1].get_theta()
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_gpu_split_lattice_adjoint_matches_factor_path():
    """Gradients through a split column's lattice solve equal the split factor
    path's to 1e-9."""
    torch = pytest.importorskip("torch")
    c = pa.SquareArray(40, 30)
    Nj, Nf, W, Nt = c.junction_count(), c.face_count(), 2, 80
    rng = np.random.default_rng(3)
    Ic = torch.tensor(np.exp(0.1 * rng.standard_normal(Nj)), dtype=torch.float64, device="cuda")
    f = torch.tensor(0.2 + 0.05 * rng.standard_normal((Nf, W)), dtype=torch.float64, device="cuda")
    Is = torch.tensor(np.outer(c.current_base(0.0), [0.5, 1.2]), dtype=torch.float64, device="cuda")
    store = np.zeros(Nt, dtype=bool)
    store[[Nt // 2, Nt - 1]] = True
    G = torch.tensor(rng.standard_normal((Nj, W, 2)), dtype=torch.float64, device="cuda")
    grads = {}
    for solver in ("lattice", "cholesky"):
        print(f"[lattice] split adjoint, {solver}", flush=True)
        with pa.DifferentiableEvolution(c, time_step=0.05, time_step_count=Nt, problem_count=W, temperature=0.01,
                                        store_time_steps=store, checkpoint_every=40, seed=9, block_size=256,
                                        face_store="global", parts=4, solver=solver) as evo:
            assert evo.eng.host.parts == 4
            assert evo.eng.host.uses_lattice == (solver == "lattice")
            leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
            (evo(*leaves) * G).sum().backward()
            grads[solver] = [t.grad.cpu().numpy() for t in leaves]
    for a, b in zip(grads["lattice"], grads["cholesky"]):
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def test_gpu_split_lattice_with_idle_parts_matches_oracle():
    """A narrow lattice (5 face columns, 3 mode pairs including the middle mode)
    split over 4 parts: the last part holds no mode pair and only meets the
    column barriers; trajectories match the oracle to 1e-9."""
    c = pa.SquareArray(30, 6)
    Nt, kw = 200, _inputs(c, W=2)
    ts = [99, Nt - 1]
    plan = CircuitPlan(c, 0.05, block_size=256, face_store="global", parts=4)
    assert plan.uses_lattice and plan.lattice.C == 5 and plan.lattice.part_local(4) == (2, 29)
    print("[lattice] split narrow lattice, 4 parts", flush=True)
    res = pa.TimeEvolutionProblem(c, time_step_count=Nt, store_time_steps=ts, solver="lattice", block_size=256,
                                  face_store="global", parts=4, temperature=0.0, **kw).compute()
    ref = to.evolve(c, dt=0.05, Nt=Nt, f=0.21, Is=kw["current_sources"], store_time_steps=ts)
    th = res.get_theta()
    assert np.abs(th - ref["theta"]).max() <= 1e-9 * np.abs(ref["theta"]).max()
    np.testing.assert_array_equal(res.get_vortex_configuration(), to.vortex(c.get_cycle_matrix(), ref["theta"]))
