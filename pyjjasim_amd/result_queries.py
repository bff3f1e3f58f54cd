"""Device-side evaluation of TimeEvolutionResult's derived quantities.

The reference answers each query with a host loop over the selected stored
steps (``_select``, time_evolution.py:1059-1067), one sparse product and, for
the phase, potential and cycle current, one SuperLU solve per step
(time_evolution.py:692-983; josephson_circuit.py:93-128). Here every selected
step of every column is one right-hand side: the stored arrays go to the device
in chunks of steps, the incidence products run as fixed-order gathers, and the
solves run as one multi-column launch of libjja's staged Cholesky solve
(jja_solve) per chunk.

Operation order follows the reference's numpy expressions (for example the
magnetic energy scales the inductance matrix by 0.5 before the product, as
``0.5 * L @ I**2`` does), so results agree with it to the rounding of the solves
and of sin/cos.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse

from .device_linalg import cut_solver, cycle_solver, _torch

# device memory one chunk of selected steps may take (bytes)
CHUNK_BYTES = 1 << 30


class PaddedRows:
    """A sparse matrix as fixed-width rows on the device: row r holds its
    column indices and coefficients in ascending column order, padded with
    column 0 and coefficient 0, so y = M x is a gather and a sum over a small
    fixed width in a fixed order."""

    def __init__(self, M, scale=1.0):
        torch = _torch()
        M = scipy.sparse.csr_matrix(M, dtype=np.float64, copy=True)
        M.sum_duplicates()
        M.sort_indices()
        r = M.shape[0]
        cnt = np.diff(M.indptr)
        k = max(1, int(cnt.max(initial=0)))
        idx = np.zeros((r, k), dtype=np.int64)
        coef = np.zeros((r, k))
        rows = np.repeat(np.arange(r), cnt)
        slot = np.arange(M.nnz) - np.repeat(M.indptr[:-1], cnt)
        idx[rows, slot] = M.indices
        coef[rows, slot] = M.data * scale if scale != 1.0 else M.data
        self.idx = torch.from_numpy(idx).cuda()
        self.coef = torch.from_numpy(coef).cuda()
        self.shape = M.shape
        self.width = k

    def __call__(self, X):
        """M @ X for X (n_cols_of_M, m) on the device."""
        return (self.coef[:, :, None] * X[self.idx]).sum(dim=1)


def _input_at(value, n, W, tps):
    """An (n, W, Nt)-broadcast input at the selected steps as an (n, W, S) array
    (callables evaluated per step, as the reference's _f/_Is do)."""
    if callable(value):
        return np.stack([np.broadcast_to(value(int(t)), (n, W)) for t in tps], axis=2)
    if hasattr(value, "amplitude") and hasattr(value, "base"):
        amp = value.amplitude[:, tps] if value.amplitude.shape[1] > 1 else value.amplitude[:, [0] * len(tps)]
        return np.broadcast_to(value.base[:, :, None] * amp[None, :, :], (n, W, len(tps)))
    return np.asarray(value)[:, :, tps]


class DeviceQueries:
    """Derived quantities of one TimeEvolutionResult, evaluated on the device."""

    def __init__(self, result):
        self.torch = _torch()
        self.res = result
        self.problem = result.problem
        self.circuit = result.problem.circuit
        self._cache = {}

    # ------------------------------------------------------------ helpers
    def _get(self, name, make):
        if name not in self._cache:
            self._cache[name] = make()
        return self._cache[name]

    def M(self):
        return self._get("M", lambda: PaddedRows(self.circuit.get_cut_matrix()))

    def A(self):
        return self._get("A", lambda: PaddedRows(self.circuit.get_cycle_matrix()))

    def L(self, scale=1.0):
        return self._get(("L", scale), lambda: PaddedRows(self.circuit._L(), scale))

    def cut(self):
        return self._get("cut", lambda: cut_solver(self.circuit))

    def cycle(self):
        return self._get("cycle", lambda: cycle_solver(self.circuit))

    def close(self):
        for v in self._cache.values():
            if hasattr(v, "close"):
                v.close()
        self._cache = {}

    def _dev(self, arr):
        return self.torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).cuda()

    def select(self, N, func, select_time_points):
        """out[:, :, i] = func(steps, stored indices) for the selected stored
        steps, evaluated in chunks of steps on the device (func returns an
        (N, W, S) device tensor)."""
        tps = np.flatnonzero(self.problem._to_time_point_mask(select_time_points))
        W = self.problem.get_problem_count()
        out = np.zeros((N, W, len(tps)))
        if len(tps) == 0:
            return out
        idx = self.res._time_point_index(tps)
        c = self.circuit
        per_step = 8 * W * (4 * max(c._Nn(), c._Nj(), c._Nf()) + 2 * c._Nj())
        step = max(1, int(CHUNK_BYTES // max(per_step, 1)))
        for s0 in range(0, len(tps), step):
            sl = slice(s0, s0 + step)
            out[:, :, sl] = func(tps[sl], idx[sl]).cpu().numpy()
        return out

    def _msq(self, X):
        """The reference's Msq_solve (last node grounded) on an (Nn, m) tensor."""
        torch = self.torch
        x = self.cut().solve(X[:-1].contiguous())
        return torch.cat((x, torch.zeros_like(X[-1:])), dim=0)

    # ------------------------------------------------------------ queries
    def phase(self, theta, tps, idx):
        Nj, W, S = theta.shape[0], theta.shape[1], len(idx)
        th = self._dev(theta[:, :, idx]).reshape(Nj, W * S)
        return self._msq(self.M()(th)).reshape(-1, W, S)

    def potential(self, voltage, tps, idx):
        return self.phase(voltage, tps, idx)

    def _series(self):
        """(terms, coefficients) of the problem's current-phase relation: 0 terms
        for the default Ic sin(theta), else its sine series (FourierCPR)."""
        from .circuit import device_cpr_terms
        series = device_cpr_terms(self.problem.current_phase_relation)
        if series is None:
            raise ValueError("device queries evaluate the default and sine-series current-phase relations")
        return series

    def josephson_energy(self, theta, tps, idx):
        """The relation's integral Ic sum_k a_k (1 - cos(k theta)) / k, in the
        operation order of FourierCPR (Ic (1 - cos theta) for the default)."""
        torch = self.torch
        th = self._dev(theta[:, :, idx])
        Ic = self._dev(self.circuit._Ic())[:, None, None]
        n, a = self._series()
        if n == 0:
            return Ic * (1.0 - torch.cos(th))
        acc = a[0] * (1.0 - torch.cos(th))
        for k in range(2, n + 1):
            acc = acc + (a[k - 1] / k) * (1.0 - torch.cos(k * th))
        return Ic * acc

    def supercurrent(self, theta, tps, idx):
        torch = self.torch
        th = self._dev(theta[:, :, idx])
        Ic = self._dev(self.circuit._Ic())[:, None, None]
        n, a = self._series()
        if n == 0:
            return Ic * torch.sin(th)
        acc = a[0] * torch.sin(th)
        for k in range(2, n + 1):
            acc = acc + a[k - 1] * torch.sin(k * th)
        return Ic * acc

    def cycle_current(self, current, tps, idx):
        Nj, W, S = current.shape[0], current.shape[1], len(idx)
        Is = self._dev(_input_at(self.problem.current_sources, Nj, W, tps))
        d = (self._dev(current[:, :, idx]) - Is).reshape(Nj, W * S)
        return self.cycle().solve(self.A()(d)).reshape(-1, W, S)

    def flux(self, current, tps, idx):
        Nj, W, S = current.shape[0], current.shape[1], len(idx)
        f = self._dev(_input_at(self.problem.external_flux, self.circuit._Nf(), W, tps))
        LI = self.L()(self._dev(current[:, :, idx]).reshape(Nj, W * S))
        return f + (self.A()(LI) / (2 * np.pi)).reshape(-1, W, S)

    def magnetic_energy(self, current, tps, idx):
        Nj, W, S = current.shape[0], current.shape[1], len(idx)
        I2 = self._dev(current[:, :, idx]).reshape(Nj, W * S) ** 2
        return self.L(0.5)(I2).reshape(Nj, W, S)

    def capacitive_energy(self, voltage, tps, idx):
        V = self._dev(voltage[:, :, idx])
        Ch = self._dev(0.5 * self.circuit._C())[:, None, None]
        return Ch * V ** 2
