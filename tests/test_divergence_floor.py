"""The floor of trajectory parity between two exact sparse solvers: the
oracle with the reference's solver (SuperLU, COLAMD order, partial pivoting,
time_evolution.py:505-506) against the oracle with a symmetric factorisation
(SuperLU in symmetric mode, minimum-degree order on K + K^T, no pivoting), on
the same inputs and noise. Both are exact to rounding; what they differ by is
what rounding alone does to the trajectory, the scale against which the
device's 1e-9 parity is measured. CPU only."""
import numpy as np
import pytest
import scipy.sparse.linalg

import pyjjasim_amd as pa
from oracle import te_oracle as to


def symmetric_solver(K):
    lu = scipy.sparse.linalg.splu(scipy.sparse.csc_matrix(K), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                                  options={"SymmetricMode": True})
    return lu.solve


def floor(c, **kw):
    a = to.evolve(c, solver=to.superlu_solver, **kw)["theta"]
    b = to.evolve(c, solver=symmetric_solver, **kw)["theta"]
    return np.abs(a - b).max() / np.abs(a).max(axis=(0, 1)).clip(1e-300)


def test_floor_plain_small_case():
    """SquareArray(10,10), f = 0.1, Is = 0.5 current_base(0), T = 0, 10,000 steps:
    the two solvers stay within 1e-11 of each other at every stored step (about
    2e-12 at the first step, where theta is smallest, 1e-13 later), so a
    1e-9 tolerance on theta is a genuine parity target over this horizon."""
    c = pa.SquareArray(10, 10)
    ts = np.array([0, 99, 999, 4999, 9999])
    f = floor(c, dt=0.05, Nt=10000, f=0.1, Is=0.5 * c.current_base(0.0)[:, None, None], store_time_steps=ts,
              store_current=False, store_voltage=False)
    print("floor by stored step:", f)
    assert f.max() < 1e-11


def test_floor_batched_iv_sweep_sample():
    """SquareArray(100,100), log-normal Ic, beta_C = 1, f = 0.25, four bias points
    of the IV sweep, T = 0.01 with the Philox stream, 300 steps."""
    c = pa.SquareArray(100, 100)
    rng = np.random.default_rng(20261202)
    c.set_critical_current(np.exp(0.1 * rng.standard_normal(c.junction_count())))
    c.set_capacitance(1.0)
    Is = c.current_base(0.0)[:, None, None] * np.array([0.25, 0.9, 1.3, 1.9])[None, :, None]
    f = floor(c, dt=0.05, Nt=300, f=0.25, Is=Is, T=0.01, store_time_steps=[99, 299], store_current=False,
              store_voltage=False, noise=to.philox_noise(12345, 4, c.junction_count()))
    print("floor by stored step:", f)
    assert f.max() < 1e-10
