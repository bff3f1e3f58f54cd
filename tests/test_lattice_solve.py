"""The lattice solve of a rectangular lattice's face system (lattice_solve.py),
on the CPU: detection of the Kronecker form, the sine tables against their
defining identities, the host emulation of the kernel's steps against a direct
sparse solve, and plan descriptors through the library's own validation."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg

import pyjjasim_amd as pa
from pyjjasim_amd import _native
from pyjjasim_amd.lattice_solve import CHUNK, LatticeSolve, lattice_form, lattice_solve
from pyjjasim_amd.solver_plan import CircuitPlan, face_system_matrix


def _xy(c):
    cx, cy = c.get_face_centroids()
    return np.stack([cx, cy], axis=1)


def _horizontal(c):
    x1, y1, x2, y2 = c.get_junction_coordinates()
    return y1 == y2


def _lattices():
    out = {"sq10": pa.SquareArray(10, 10), "sq100": pa.SquareArray(100, 100), "wide": pa.SquareArray(17, 6),
           "tall": pa.SquareArray(5, 23), "sq3": pa.SquareArray(3, 3)}
    c = pa.SquareArray(12, 9)
    c.set_inductance(0.7)
    c.set_capacitance(1.3)
    out["selfL"] = c
    c = pa.SquareArray(9, 14)
    c.set_capacitance(np.where(_horizontal(c), 1.0, 2.5))     # one weight per direction
    out["aniso"] = c
    return out


def _not_lattices():
    out = {"hc": pa.HoneycombArray(6, 5), "tri": pa.TriangularArray(6, 6), "squid": pa.SQUID(),
           "strip": pa.SquareArray(8, 2)}
    c = pa.SquareArray(8, 8)
    R = np.ones(c.junction_count())
    R[5] = 1.5
    c.set_resistance(R)
    out["hetero_R"] = c
    c = pa.SquareArray(6, 6)
    Nj = c.junction_count()
    L = scipy.sparse.diags(0.5 * np.ones(Nj)) + scipy.sparse.diags(0.1 * np.ones(Nj - 1), 1) + \
        scipy.sparse.diags(0.1 * np.ones(Nj - 1), -1)
    c.set_inductance(L)
    out["mutual_L"] = c
    return out


@pytest.mark.parametrize("name", sorted(_lattices()))
def test_lattice_solve_matches_direct_solve(name):
    c = _lattices()[name]
    K = face_system_matrix(c, 0.05)
    lat = lattice_solve(K, _xy(c))
    assert lat is not None and lat.C <= lat.R and lat.R * lat.C == K.shape[0]
    b = np.random.default_rng(3).standard_normal((K.shape[0], 3))
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(K), b)
    got = lat.emulate(b)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("name", sorted(_not_lattices()))
def test_lattice_form_rejects_other_face_systems(name):
    c = _not_lattices()[name]
    assert lattice_form(face_system_matrix(c, 0.05), _xy(c)) is None
    with pytest.raises(ValueError):
        CircuitPlan(c, 0.05, solver="lattice")


@pytest.mark.parametrize("C_", [2, 3, 9, 14, 99, 100])
def test_sine_tables_are_the_orthonormal_dst(C_):
    """S S = I, S T_C S = diag(mu) from the tabulated half of S unfolded with
    S[C-1-i][j] = (-1)^j S[i][j]; the zero padding the kernel reads is zero."""
    lat = LatticeSolve(4, C_, np.arange(4 * C_), 1.0, 1.0)
    H = lat.H
    S = np.zeros((C_, C_))
    S[:H] = lat.sine[:H, :C_]
    for i in range(H, C_):
        S[i] = lat.sine[C_ - 1 - i, :C_] * (-1.0) ** np.arange(C_)
    np.testing.assert_allclose(S @ S, np.eye(C_), atol=1e-14)
    T = 2 * np.eye(C_) - np.eye(C_, k=1) - np.eye(C_, k=-1)
    mu = 4 * np.sin(np.pi * np.arange(1, C_ + 1) / (2 * (C_ + 1))) ** 2
    np.testing.assert_allclose(S @ T @ S, np.diag(mu), atol=1e-13)
    assert lat.CP % CHUNK == 0 and lat.ES % 2 == 0 and lat.ES >= H
    assert not lat.sine[H:].any() and not lat.sine[:, C_:].any()


def test_thomas_pivots_factor_each_mode():
    """inv[r][k] are the inverse pivots of g T_R + g_c mu_k I: the factor
    L U they define reproduces the block."""
    R, C_, g, gc = 7, 5, 1.7, 0.6
    lat = LatticeSolve(R, C_, np.arange(R * C_), g, gc)
    mu = 4 * np.sin(np.pi * np.arange(1, C_ + 1) / (2 * (C_ + 1))) ** 2
    T = 2 * np.eye(R) - np.eye(R, k=1) - np.eye(R, k=-1)
    for k in range(C_):
        M = g * T + gc * mu[k] * np.eye(R)
        piv = 1.0 / lat.inv[:, k]
        Lo = np.eye(R) + np.diag(-g / piv[:-1], -1)
        U = np.diag(piv) + np.diag(-g * np.ones(R - 1), 1)
        np.testing.assert_allclose(Lo @ U, M, atol=1e-13)


@pytest.mark.parametrize("store", ["lds", "global"])
def test_lattice_plan_descriptor_passes_native_validation(store):
    c = pa.SquareArray(12, 9)
    plan = CircuitPlan(c, 0.05, block_size=256, face_store=store)
    assert plan.uses_lattice
    assert plan.index_bytes == (4 if store == "global" else 2)
    d = plan.desc()
    assert d.lattice_rows * d.lattice_cols == c.face_count() and d.n_solve_slices == 0
    _native.check(_native.lib().jja_plan_validate(1, C.byref(d)))
    # the face phase by lattice position, then the lattice solve, equals the factor program
    u = np.random.default_rng(5).standard_normal(c.junction_count())
    ref = CircuitPlan(c, 0.05, block_size=256, face_store=store, solver="cholesky")
    assert not ref.uses_lattice
    want = ref.emulate_face_and_solve(u)
    np.testing.assert_allclose(plan.emulate_face_and_solve(u), want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_lattice_descriptor_errors_are_refused():
    c = pa.SquareArray(8, 7)
    plan = CircuitPlan(c, 0.05, block_size=128)
    lib = _native.lib()
    d = plan.desc()
    d.max_stage_rows = d.max_stage_rows - 2
    assert lib.jja_plan_validate(1, C.byref(d)) == _native.JJA_EINVAL
    d = plan.desc()
    d.lattice_cols = d.lattice_cols + 1
    assert lib.jja_plan_validate(1, C.byref(d)) == _native.JJA_EINVAL
    d = plan.desc()
    d.lattice_g = float("nan")
    assert lib.jja_plan_validate(1, C.byref(d)) == _native.JJA_EINVAL
    # a split column's tridiagonal solves take more blocks, so more scratch
    split = CircuitPlan(c, 0.05, block_size=128, parts=2)
    assert split.uses_lattice and split.global_store
    assert lib.jja_plan_validate(1, C.byref(split.desc())) == _native.JJA_OK
    wide = pa.SquareArray(200, 20)
    one, four = (CircuitPlan(wide, 0.05, block_size=1024, parts=p) for p in (1, 4))
    assert four.lattice.tri_blocks(4) > one.lattice.tri_blocks(1)
    d = four.desc()
    d.max_stage_rows = one.relayout(1024, face_store="global").scratch_rows
    assert d.max_stage_rows < four.scratch_rows
    assert lib.jja_plan_validate(1, C.byref(d)) == _native.JJA_EINVAL


def test_solver_choice():
    c = pa.SquareArray(10, 8)
    assert CircuitPlan(c, 0.05).uses_lattice
    assert not CircuitPlan(c, 0.05, solver="cholesky").uses_lattice
    p = CircuitPlan(c, 0.05, solver="cholesky").relayout(1024, solver="auto")
    assert p.uses_lattice
    sp = p.relayout(1024, parts=2, face_store="global")
    assert sp.uses_lattice and sp.parts == 2
    # column pairs share factor records, which a lattice plan does not have
    with pytest.raises(ValueError):
        sp.paired()
    assert not sp.relayout(1024, solver="cholesky").paired().uses_lattice
    with pytest.raises(ValueError):
        CircuitPlan(pa.HoneycombArray(4, 4), 0.05, solver="lattice")
    with pytest.raises(ValueError):
        CircuitPlan(c, 0.05, solver="fft")


@pytest.mark.parametrize("parts", [2, 4])
def test_split_lattice_emulation_matches_the_factor(parts):
    """A split column cuts each mode's recurrence into more blocks: the solve
    still equals the factor's to rounding."""
    c = pa.SquareArray(200, 20)
    plan = CircuitPlan(c, 0.05, block_size=256, parts=parts)
    assert plan.uses_lattice and plan.lattice.tri_blocks(parts) > plan.lattice.tri_blocks(1)
    ref = CircuitPlan(c, 0.05, block_size=256, solver="cholesky")
    u = np.random.default_rng(parts).standard_normal(c.junction_count())
    want = ref.emulate_face_and_solve(u)
    np.testing.assert_allclose(plan.emulate_face_and_solve(u), want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_split_balancing_prices_lattice_plans_by_their_transforms():
    """A split launch keeps the lattice solve, so the part balancing prices a
    lattice problem by its transforms and face passes (less than its factor
    for a large lattice) and gives it no more parts than the factor would."""
    from pyjjasim_amd.batch import _split_cost, balanced_parts
    c = pa.SquareArray(120, 120)
    lat = CircuitPlan(c, 0.05, block_size=1024)
    chol = CircuitPlan(c, 0.05, block_size=1024, solver="cholesky")
    assert lat.uses_lattice and not chol.uses_lattice
    assert 8 * lat.Nj < _split_cost(lat) < _split_cost(chol)
    hc = CircuitPlan(pa.HoneycombArray(60, 60), 0.05, block_size=1024)
    p_lat, p_chol = (balanced_parts([h, hc], [1, 1], 16) for h in (lat, chol))
    assert p_lat[0] <= p_chol[0] and p_lat[1] >= p_chol[1]
