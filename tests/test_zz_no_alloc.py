"""jja_solve_noalloc (the device scratch variant of jja_solve) through the C
ABI: on a matrix plan with the global face store, two calls in a row, the
second with a smaller input, each equal bit for bit to jja_solve with a
caller workspace; a call needing more than the scratch is refused."""
import ctypes as C

import numpy as np
import pytest

import pyjjasim_amd as pa
from pyjjasim_amd import _native

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _solver(n):
    from pyjjasim_amd.device_linalg import DeviceSolver
    c = pa.SquareArray(n, n)
    A = c.get_cycle_matrix()
    xy = np.stack(c.get_face_centroids(), axis=1)
    return DeviceSolver((A @ A.T).tocsc(), xy, face_store="global", max_columns=64)


def test_gpu_solve_noalloc_equals_solve():
    print("[no-alloc] jja_solve_noalloc vs jja_solve", flush=True)
    rng = np.random.default_rng(0)
    for n, cols in ((60, 12), (30, 3)):               # the second call is smaller
        s = _solver(n)
        try:
            assert s.zd > 0                          # the global face store is in use
            B = torch.tensor(rng.standard_normal((s.n, cols)), device="cuda")
            want = s.solve(B)
            got = s.solve_noalloc(B)
            torch.cuda.synchronize()
            assert torch.equal(got, want)
            K = s.host.K
            res = np.abs(K @ got.cpu().numpy() - B.cpu().numpy()).max()
            assert res < 1e-10 * np.abs(B.cpu().numpy()).max()
        finally:
            s.close()


def test_gpu_solve_noalloc_refuses_beyond_scratch():
    s = _solver(60)
    try:
        cols = (1 << 21) // s.zd + 1                  # one column more than the scratch holds
        b = torch.zeros((cols, s.n), dtype=torch.float64, device="cuda")
        x = torch.empty_like(b)
        rc = s.lib.jja_solve_noalloc(s.handle, 0, b.data_ptr(), s.n, x.data_ptr(), s.n, cols,
                                     torch.cuda.current_stream().cuda_stream)
        assert rc == _native.JJA_EINVAL
        assert b"scratch" in s.lib.jja_last_error()
    finally:
        s.close()
