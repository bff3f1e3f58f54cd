"""Failure handling on the device path, each condition forced from the host so
that every kernel still runs to its end: a column barrier that gives up, a
split launch larger than the device holds at once, and a run missing its
global face store."""
import numpy as np
import pytest

import pyjjasim_amd as pa
from pyjjasim_amd import _native
from pyjjasim_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _problem_f(*args, **kwargs):
    """A problem solved with the factor program (the subject here), also where
    the lattice solve would take it."""
    kwargs.setdefault("solver", "cholesky")
    return pa.TimeEvolutionProblem(*args, **kwargs)


def _evolution_f(*args, **kwargs):
    kwargs.setdefault("solver", "cholesky")
    return pa.DifferentiableEvolution(*args, **kwargs)


def _problem(W=2, Nt=40, parts=4, **kw):
    c = pa.SquareArray(12, 10)
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.5, 1.2, W)[None, :, None]
    return _problem_f(c, time_step=0.05, time_step_count=Nt, external_flux=0.1,
                                   current_sources=Is, store_time_steps=[Nt - 1], block_size=256,
                                   face_store="global", parts=parts, **kw)


def _masks(p):
    return p.store_time_steps.copy(), np.zeros(p._Nt(), dtype=bool)


def test_gpu_barrier_give_up_is_reported():
    """With a wait limit of zero polls every column barrier gives up at once:
    the launch still completes, and jja_plan_status (called by results())
    reports the invalid run; after that, a run with the default limit is exact."""
    print("[errors] barrier give-up", flush=True)
    p = _problem()
    eng = Engine(p, *_masks(p))
    try:
        _native.check(eng.lib.jja_plan_set_wait_limit(eng.handle, 0))
        eng.advance(p._Nt())
        with pytest.raises(_native.NativeError, match="gave up"):
            eng.results()
    finally:
        eng.close()
    good = _problem().compute()
    ref = _problem(parts=1).compute()
    np.testing.assert_array_equal(good.get_theta(), ref.get_theta())


def test_gpu_split_launch_larger_than_the_device_is_refused():
    """W x parts workgroups beyond what the device holds at once: refused with
    an error before anything runs (a cooperative launch cannot wait for room)."""
    print("[errors] oversized split launch", flush=True)
    import torch
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    W = 8 * n_cu                     # 8 W x 4 parts x 256 threads: beyond any residency
    with pytest.raises(_native.NativeError, match="split launch"):
        _problem(W=W, Nt=4).compute()


def test_gpu_missing_face_store_is_rejected():
    """A run of a global-face-store problem without its zbuf array is refused
    (ValueError) before any launch."""
    print("[errors] missing zbuf", flush=True)
    p = _problem(parts=1)
    eng = Engine(p, *_masks(p))
    try:
        assert eng.zbuf is not None
        eng.run.zbuf = None
        with pytest.raises(ValueError, match="zbuf"):
            eng.launch(1)
        eng.run.zbuf = eng.zbuf.data_ptr()
        eng.launch(1)
    finally:
        eng.close()


def test_gpu_plan_busy_on_another_stream_is_refused():
    """The plan's launches share its device descriptors, so they run on one
    stream at a time: while a long launch runs on stream A, a launch on stream
    B is refused (JJA_EBUSY) before anything is enqueued; once A has finished,
    B may launch, and the run equals one made on a single stream."""
    print("[errors] busy plan on another stream", flush=True)
    import torch

    def problem():
        c = pa.SquareArray(60, 60)
        Is = c.current_base(0.0)[:, None, None] * np.linspace(0.5, 1.5, 64)[None, :, None]
        return _problem_f(c, time_step=0.05, time_step_count=4001, external_flux=0.2,
                                       current_sources=Is, store_time_steps=[4000], store_voltage=False,
                                       store_current=False)
    p = problem()
    eng = Engine(p, *_masks(p))
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    try:
        eng.launch(4000, stream=sa.cuda_stream)
        with pytest.raises(_native.NativeError, match="another stream"):
            eng.launch(1, stream=sb.cuda_stream)
        sa.synchronize()
        eng.launch(1, stream=sb.cuda_stream)
        sb.synchronize()
        th = eng.results()[0]
    finally:
        eng.close()
    ref = problem().compute()
    np.testing.assert_array_equal(th[:, :, -1], ref.get_theta()[:, :, -1])


def test_gpu_more_columns_than_planned_is_refused():
    """A run with more columns than the plan was created for (its launch
    descriptors are sized at creation) is refused with ValueError."""
    print("[errors] more columns than planned", flush=True)
    p = _problem(parts=1)
    eng = Engine(p, *_masks(p))
    try:
        eng.run.W = p.problem_count + 1
        with pytest.raises(ValueError, match="max_columns"):
            eng.launch(1)
    finally:
        eng.close()


def test_gpu_split_differentiable_evolution_reports_give_up():
    """A split DifferentiableEvolution whose column barriers give up (wait limit
    zero, every kernel still running to its end) raises NativeError from the
    forward pass instead of returning invalid phases; with the default limit
    its forward and gradient equal those of one workgroup per column."""
    print("[errors] split DifferentiableEvolution give-up", flush=True)
    import torch
    c = pa.SquareArray(12, 10)
    W, Nt = 2, 40
    Nj, Nf = c.junction_count(), c.face_count()
    Ic = torch.ones(Nj, dtype=torch.float64, device="cuda")
    f = torch.full((Nf,), 0.1, dtype=torch.float64, device="cuda")
    Is = torch.tensor(c.current_base(0.0)[:, None] * np.array([0.6, 1.3])[None, :], device="cuda")
    grads = []
    for parts in (4, 1):
        with _evolution_f(c, time_step=0.05, time_step_count=Nt, problem_count=W,
                                        store_time_steps=[Nt - 1], checkpoint_every=16, block_size=256,
                                        face_store="global", parts=parts) as evo:
            if parts > 1:
                _native.check(evo.eng.lib.jja_plan_set_wait_limit(evo.eng.handle, 0))
                with pytest.raises(_native.NativeError, match="gave up"):
                    evo(Ic, f, Is)
                _native.check(evo.eng.lib.jja_plan_set_wait_limit(evo.eng.handle, 1 << 21))
            leaf = Ic.clone().requires_grad_(True)
            th = evo(leaf, f, Is)
            th.sum().backward()
            grads.append((th.detach().cpu().numpy(), leaf.grad.cpu().numpy()))
    np.testing.assert_array_equal(grads[0][0], grads[1][0])
    np.testing.assert_array_equal(grads[0][1], grads[1][1])


def test_gpu_plans_with_different_lds_interleave():
    """The dynamic LDS limit belongs to the kernel, not to a plan: launching a
    plan that needs little LDS between two launches of one that needs 157 KB
    must not lower the limit under the larger plan."""
    def prob(n, Nt=30):
        c = pa.SquareArray(n, n)
        Is = c.current_base(0.0)[:, None, None] * np.array([0.5, 1.1])[None, :, None]
        return pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=Nt, external_flux=0.2,
                                       current_sources=Is, block_size=1024, store_time_steps=[Nt - 1])
    print("[errors] interleaved plans with different LDS", flush=True)
    big, small = prob(100), prob(6)
    eb = pa.TimeEvolutionBatch([big], block_size=1024).make_engine()
    es = pa.TimeEvolutionBatch([small], block_size=1024).make_engine()
    try:
        eb.advance(10)
        es.advance(10)
        eb.advance(20)
        es.advance(20)
        th_big = eb.results()[0][0]
        th_small = es.results()[0][0]
    finally:
        eb.close()
        es.close()
    np.testing.assert_array_equal(th_big[:, :, -1], big.compute().get_theta()[:, :, -1])
    np.testing.assert_array_equal(th_small[:, :, -1], small.compute().get_theta()[:, :, -1])
