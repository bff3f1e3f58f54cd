"""Device linear algebra for result queries, static solves and annealing: sparse
SPD solves with many right-hand sides on the GPU (libjja's staged Cholesky
solve program, jja_solve) and signed incidence products as torch sparse CSR
matrices.

The reference answers these with scipy on the host: ``factorized`` for the
cycle system A A^T and the grounded cut system Mr Mr^T
(josephson_circuit.py:93-128) and sparse ``@`` with A and M. Here the factor is
computed once on the host (native Cholesky in nested-dissection order) and
every solve runs on the device.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import scipy.sparse

from . import _native
from .solver_plan import MatrixPlan


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pyjjasim_amd: device linear algebra needs a HIP (ROCm) device; none is visible")
    return torch


class DeviceSolver:
    """K x = b on the device for a sparse symmetric positive definite K
    (factorised once). ``solve(B)``: B a float64 device tensor (n,) or (n, k);
    returns a new tensor of the same shape."""

    def __init__(self, K, xy, leaf_size=16, block_size=None, max_columns=1024, face_store="auto"):
        self.torch = torch = _torch()
        self.lib = _native.lib()
        self.host = MatrixPlan(K, xy, leaf_size=leaf_size, block_size=block_size, face_store=face_store)
        self.n = self.host.Nf
        self._desc = self.host.desc(max_columns=max_columns)
        self.handle = C.c_void_p()
        _native.check(self.lib.jja_plan_create(1, C.byref(self._desc), C.byref(self.handle)))
        self.zd = int(self.lib.jja_plan_zbuf_doubles(self.handle, 0))
        if self.zd == 0 and self.lib.jja_plan_lds_bytes(self.handle, 0) == 0:
            self.close()
            raise _native.NativeError(f"matrix of order {self.n} needs more LDS per column than a workgroup has")
        self._zbuf = None

    def solve(self, B):
        torch = self.torch
        squeeze = B.dim() == 1
        Bc = (B[:, None] if squeeze else B).t().contiguous()      # [cols][n]
        if Bc.dtype != torch.float64 or Bc.shape[1] != self.n:
            raise ValueError(f"right-hand sides must be float64 with {self.n} rows")
        cols = Bc.shape[0]
        X = torch.empty_like(Bc)
        zb = None
        if self.zd:
            if self._zbuf is None or self._zbuf.shape[0] < cols:
                self._zbuf = torch.empty((cols, self.zd), dtype=torch.float64, device=Bc.device)
            zb = self._zbuf.data_ptr()
        stream = torch.cuda.current_stream().cuda_stream
        _native.check(self.lib.jja_solve(self.handle, 0, Bc.data_ptr(), self.n, X.data_ptr(), self.n, cols,
                                         zb, stream))
        X = X.t()
        return X[:, 0] if squeeze else X

    def solve_noalloc(self, B):
        """solve() through jja_solve_noalloc: the face store comes from the
        library's fixed device scratch (16 MiB per GPU) instead of a workspace,
        and nothing is allocated below this call except the output tensor."""
        torch = self.torch
        squeeze = B.dim() == 1
        Bc = (B[:, None] if squeeze else B).t().contiguous()
        if Bc.dtype != torch.float64 or Bc.shape[1] != self.n:
            raise ValueError(f"right-hand sides must be float64 with {self.n} rows")
        X = torch.empty_like(Bc)
        stream = torch.cuda.current_stream().cuda_stream
        _native.check(self.lib.jja_solve_noalloc(self.handle, 0, Bc.data_ptr(), self.n, X.data_ptr(), self.n,
                                                 Bc.shape[0], stream))
        X = X.t()
        return X[:, 0] if squeeze else X

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.jja_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        # at interpreter exit the HIP runtime's own exit handlers may already
        # have run: leave the memory to the process teardown instead of calling
        # into the runtime (close() explicitly to free it earlier)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def csr_tensor(M, device="cuda"):
    """A scipy sparse matrix as a float64 torch sparse CSR tensor on the device."""
    torch = _torch()
    M = scipy.sparse.csr_matrix(M, dtype=np.float64)
    M.sort_indices()
    return torch.sparse_csr_tensor(torch.from_numpy(M.indptr.astype(np.int64)),
                                   torch.from_numpy(M.indices.astype(np.int64)),
                                   torch.from_numpy(M.data), size=M.shape, dtype=torch.float64,
                                   device=device)


def cycle_solver(circuit):
    """Solver of A A^T (the reference's Asq_solve, josephson_circuit.py:111-128)."""
    A = circuit.get_cycle_matrix()
    cx, cy = circuit.get_face_centroids()
    return DeviceSolver((A @ A.T).tocsc(), np.stack([cx, cy], axis=1))


def cut_solver(circuit):
    """Solver of Mr Mr^T, the cut-space system with the last node grounded (the
    reference's Msq_solve, josephson_circuit.py:93-109)."""
    Mr = circuit._Mr()
    x, y = circuit.get_node_coordinates()
    return DeviceSolver((Mr @ Mr.T).tocsc(), np.stack([x[:-1], y[:-1]], axis=1))
