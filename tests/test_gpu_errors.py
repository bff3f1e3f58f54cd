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


def _problem(W=2, Nt=40, parts=4, **kw):
    c = pa.SquareArray(12, 10)
    Is = c.current_base(0.0)[:, None, None] * np.linspace(0.5, 1.2, W)[None, :, None]
    return pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=Nt, external_flux=0.1,
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
