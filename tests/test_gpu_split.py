"""A column split over several workgroups (parts) that meet at column barriers
in global memory gives exactly the results of one workgroup per column, and so
does a solve whose lowest stages run wave-locally: the per-row summation order,
the Philox stream and every expression are the same."""
import numpy as np
import pytest

import pyjjasim_amd as pa

pytestmark = pytest.mark.gpu


def _problem_f(*args, **kwargs):
    """A problem solved with the factor program (the subject here), also where
    the lattice solve would take it."""
    kwargs.setdefault("solver", "cholesky")
    return pa.TimeEvolutionProblem(*args, **kwargs)


def _evolution_f(*args, **kwargs):
    kwargs.setdefault("solver", "cholesky")
    return pa.DifferentiableEvolution(*args, **kwargs)


def _circuit():
    c = pa.SquareArray(36, 28)
    rng = np.random.default_rng(21)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    c.set_capacitance(0.8)
    c.set_inductance(0.3)
    return c


def _run(c, parts, T, Nt=150, store="global", local="auto"):
    W = 3
    Is = c.current_base(0.0)[:, None, None] * np.array([0.4, 0.9, 1.3])[None, :, None]
    prob = _problem_f(c, time_step=0.05, time_step_count=Nt, external_flux=0.17,
                                   current_sources=Is, temperature=T, seed=5,
                                   store_time_steps=[Nt // 2, Nt - 1], block_size=256,
                                   face_store=store, parts=parts, steps_per_launch=60,
                                   local_stages=local)
    return prob.compute()


@pytest.mark.parametrize("T", [0.0, 0.02])
def test_gpu_split_columns_bit_identical(T):
    c = _circuit()
    print(f"[split] one part, T={T}", flush=True)
    ref = _run(c, 1, T)
    for parts in (2, 4):
        print(f"[split] {parts} parts, T={T}", flush=True)
        r = _run(c, parts, T)
        np.testing.assert_array_equal(r.get_theta(), ref.get_theta())
        np.testing.assert_array_equal(r.get_voltage(), ref.get_voltage())
        np.testing.assert_array_equal(r.get_vortex_configuration(), ref.get_vortex_configuration())


def _engine_run(c, parts, W, Nt=150):
    """theta output and the count of workgroups whose column barriers found
    every part of their column on one XCD, over the run's launches."""
    from pyjjasim_amd.engine import Engine
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.3, 1.4, W)[None, :, None]
    prob = _problem_f(c, time_step=0.05, time_step_count=Nt, external_flux=0.17,
                                   current_sources=Is, temperature=0.02, seed=9,
                                   store_time_steps=[Nt // 2, Nt - 1], block_size=256,
                                   face_store="global", parts=parts, steps_per_launch=50)
    eng = Engine(prob, prob.store_time_steps.copy(), np.zeros(Nt, dtype=bool))
    try:
        eng.advance(Nt)
        th, _, nv = eng.results()
        n = eng.lib.jja_plan_one_xcd_parts(eng.handle)
        assert n >= 0
    finally:
        eng.close()
    return th, nv, n


@pytest.mark.parametrize("W,parts,packed", [(8, 2, True), (8, 4, True), (16, 4, True), (3, 4, False)])
def test_gpu_split_one_xcd_barriers_bit_identical(W, parts, packed):
    """Launches whose workgroups divide evenly over the 8 XCDs place every
    column's parts on one XCD (block indices equal modulo 8); their column
    barriers then skip the L2 write-back. W=3 x 4 parts cannot
    be packed, and its barriers take the device-scope path. Either way the
    results are those of one workgroup per column."""
    c = _circuit()
    print(f"[xcd] W={W}, one part", flush=True)
    ref_th, ref_nv, n1 = _engine_run(c, 1, W)
    assert n1 == 0
    print(f"[xcd] W={W}, {parts} parts", flush=True)
    th, nv, n = _engine_run(c, parts, W)
    np.testing.assert_array_equal(th, ref_th)
    np.testing.assert_array_equal(nv, ref_nv)
    launches = 3
    assert n == (launches * W * parts if packed else 0)


@pytest.mark.parametrize("store,parts,local", [("lds", 1, 3), ("lds", 1, 99), ("global", 1, 99),
                                               ("global", 4, 0), ("global", 4, 2)])
def test_gpu_wave_local_stages_bit_identical(store, parts, local):
    """Lowest stages solved wave-locally (each wave whole subtrees, no barrier
    between those stages), in LDS, in global memory and split over
    workgroups, give the results of every stage solved in step."""
    c = _circuit()
    print("[local] reference: one part, every stage in step", flush=True)
    ref = _run(c, 1, 0.02, local=0)
    print(f"[local] {store} store, {parts} parts, local_stages={local}", flush=True)
    r = _run(c, parts, 0.02, store=store, local=local)
    np.testing.assert_array_equal(r.get_theta(), ref.get_theta())
    np.testing.assert_array_equal(r.get_voltage(), ref.get_voltage())
    np.testing.assert_array_equal(r.get_vortex_configuration(), ref.get_vortex_configuration())


def _run_pairs(c, W, pairs, Nt=120):
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.3, 1.4, W)[None, :, None]
    prob = _problem_f(c, time_step=0.05, time_step_count=Nt, external_flux=0.17,
                                   current_sources=Is, temperature=0.02, seed=11,
                                   store_time_steps=[Nt // 2, Nt - 1], block_size=512,
                                   face_store="global", parts=2, steps_per_launch=50,
                                   column_pairs=pairs)
    return prob.compute()


@pytest.mark.parametrize("W", [2, 3, 8])
def test_gpu_column_pairs_bit_identical(W):
    """Split columns worked two at a time (one record stream, one barrier for
    both; the last column alone when W is odd) give the results of unpaired
    split columns."""
    c = _circuit()
    print(f"[pairs] W={W}, unpaired", flush=True)
    ref = _run_pairs(c, W, False)
    print(f"[pairs] W={W}, paired", flush=True)
    r = _run_pairs(c, W, True)
    np.testing.assert_array_equal(r.get_theta(), ref.get_theta())
    np.testing.assert_array_equal(r.get_voltage(), ref.get_voltage())
    np.testing.assert_array_equal(r.get_vortex_configuration(), ref.get_vortex_configuration())


def test_gpu_split_adjoint_bit_identical():
    torch = pytest.importorskip("torch")
    c = _circuit()
    Nj, Nf, W, Nt = c.junction_count(), c.face_count(), 3, 80
    rng = np.random.default_rng(8)
    Ic = torch.tensor(np.exp(0.2 * rng.standard_normal(Nj)), dtype=torch.float64, device="cuda")
    f = torch.tensor(0.1 + 0.02 * rng.standard_normal(Nf), dtype=torch.float64, device="cuda")
    Is = torch.tensor(c.current_base(0.0)[:, None] * np.linspace(0.5, 1.5, W)[None, :],
                      dtype=torch.float64, device="cuda")
    store = np.zeros(Nt, dtype=bool)
    store[[Nt // 2, Nt - 1]] = True
    G = torch.tensor(rng.standard_normal((Nj, W, 2)), dtype=torch.float64, device="cuda")
    grads = []
    # one part in step, four parts (wave-local lowest stages by default), one
    # part with every stage wave-local, and columns paired (512-thread
    # workgroups halved, 2 parts doubled)
    for parts, local, bs, pairs in ((1, 0, 256, False), (4, "auto", 256, False), (1, 99, 256, False),
                                    (2, "auto", 512, True)):
        print(f"[split] adjoint, {parts} parts, local_stages={local}, pairs={pairs}", flush=True)
        evo = _evolution_f(c, time_step=0.05, time_step_count=Nt, problem_count=W,
                                         temperature=0.01, store_time_steps=store, checkpoint_every=30,
                                         block_size=bs, face_store="global", parts=parts, seed=3,
                                         local_stages=local, column_pairs=pairs)
        leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
        th = evo(*leaves)
        (th * G).sum().backward()
        grads.append([th.detach().cpu().numpy()] + [t.grad.cpu().numpy() for t in leaves])
    for other in grads[1:]:
        for a, b in zip(grads[0], other):
            np.testing.assert_array_equal(a, b)


def test_gpu_part_local_stages_bit_identical():
    """A split column whose middle factor stages are part-local (each workgroup
    solves its own subtrees between workgroup barriers) gives the results of
    one workgroup per column and of every stage solved in step."""
    from pyjjasim_amd.solver_plan import CircuitPlan
    c = pa.SquareArray(64, 64)
    c.set_capacitance(0.9)
    assert CircuitPlan(c, 0.05, block_size=512, parts=2, solver="cholesky").part_stages > 0
    W, Nt = 2, 60
    Is = c.current_base(0.0)[:, None, None] * np.array([0.6, 1.2])[None, :, None]

    def run(parts, local):
        prob = _problem_f(c, time_step=0.05, time_step_count=Nt, external_flux=0.13,
                                       current_sources=Is, temperature=0.02, seed=13,
                                       store_time_steps=[Nt // 2, Nt - 1], block_size=512,
                                       face_store="global", parts=parts, steps_per_launch=25,
                                       local_stages=local)
        return prob.compute()
    print("[part-local] one part", flush=True)
    ref = run(1, 0)
    for parts, local in ((2, "auto"), (2, 0)):
        print(f"[part-local] {parts} parts, local_stages={local}", flush=True)
        r = run(parts, local)
        np.testing.assert_array_equal(r.get_theta(), ref.get_theta())
        np.testing.assert_array_equal(r.get_vortex_configuration(), ref.get_vortex_configuration())


def test_gpu_part_local_adjoint_bit_identical():
    """The adjoint of a split column with part-local stages (the same programs,
    workgroup barriers included, swept backwards) gives the gradients of one
    workgroup per column."""
    torch = pytest.importorskip("torch")
    c = pa.SquareArray(64, 64)
    Nj, Nf, W, Nt = c.junction_count(), c.face_count(), 2, 40
    rng = np.random.default_rng(17)
    Ic = torch.tensor(np.exp(0.2 * rng.standard_normal(Nj)), dtype=torch.float64, device="cuda")
    f = torch.tensor(0.1 + 0.02 * rng.standard_normal(Nf), dtype=torch.float64, device="cuda")
    Is = torch.tensor(c.current_base(0.0)[:, None] * np.array([0.7, 1.3])[None, :],
                      dtype=torch.float64, device="cuda")
    store = np.zeros(Nt, dtype=bool)
    store[[Nt // 2, Nt - 1]] = True
    G = torch.tensor(rng.standard_normal((Nj, W, 2)), dtype=torch.float64, device="cuda")
    grads = []
    for parts in (1, 2):
        print(f"[part-local] adjoint, {parts} parts", flush=True)
        evo = _evolution_f(c, time_step=0.05, time_step_count=Nt, problem_count=W,
                                         temperature=0.01, store_time_steps=store, checkpoint_every=20,
                                         block_size=512, face_store="global", parts=parts, seed=5)
        if parts == 2:
            assert evo.eng.host.part_stages > 0
        leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
        th = evo(*leaves)
        (th * G).sum().backward()
        grads.append([th.detach().cpu().numpy()] + [t.grad.cpu().numpy() for t in leaves])
    for a, b in zip(*grads):
        np.testing.assert_array_equal(a, b)
