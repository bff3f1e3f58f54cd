"""Host preparation for the GPU: native Cholesky, nested-dissection stage schedule,
cycle-matrix index arrays. The schedule is checked by emulating on the host
exactly the stage algorithm the kernel runs, against a direct sparse solve."""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg

import pyjjasim_amd as pa
from pyjjasim_amd.solver_plan import CircuitPlan, cholesky, cycle_codes


def _plan(*args, **kwargs):
    """A plan with the staged factor program (the subject of these tests), also
    for lattices the lattice solve would take."""
    kwargs.setdefault("solver", "cholesky")
    return CircuitPlan(*args, **kwargs)


def _circuits():
    sqL = pa.SquareArray(12, 9)
    sqL.set_inductance(0.8)
    het = pa.SquareArray(9, 9)
    rng = np.random.default_rng(1)
    het.set_resistance(0.5 + rng.random(het.junction_count()))
    het.set_capacitance(rng.random(het.junction_count()))
    return {"sq10": pa.SquareArray(10, 10), "hc7": pa.HoneycombArray(7, 5),
            "tri8": pa.TriangularArray(8, 8), "squid": pa.SQUID(), "sqL": sqL, "het": het}


@pytest.mark.parametrize("name", sorted(_circuits()))
def test_staged_solve_matches_direct(name):
    c = _circuits()[name]
    plan = _plan(c, 0.05, leaf_size=8)
    K = scipy.sparse.csc_matrix(plan.K)
    b = np.random.default_rng(0).standard_normal((K.shape[0], 4))
    want = scipy.sparse.linalg.spsolve(K, b)
    got = plan.solve_plan.solve(b)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * abs(want).max())


def test_cholesky_reproduces_matrix():
    rng = np.random.default_rng(5)
    n = 60
    B = scipy.sparse.random(n, n, density=0.08, random_state=7)
    K = (B @ B.T + scipy.sparse.identity(n) * 3).tocsc()
    L = cholesky(K)
    assert abs(L @ L.T - K).max() < 1e-12
    assert abs(scipy.sparse.triu(L, 1)).max() == 0 if scipy.sparse.triu(L, 1).nnz else True
    np.testing.assert_allclose(L.toarray(), np.linalg.cholesky(K.toarray()), atol=1e-12)


def test_cholesky_rejects_indefinite():
    K = scipy.sparse.csc_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))
    with pytest.raises(ValueError, match="positive definite"):
        cholesky(K)


def test_schedule_structure():
    c = pa.SquareArray(30, 30)
    sp = _plan(c, 0.05, leaf_size=16).solve_plan
    n = sp.n
    assert sp.stage_ptr[0] == 0 and sp.stage_ptr[-1] == n
    # off-block rows only reach earlier stages (lower) or later stages (upper)
    stage_of = np.searchsorted(sp.stage_ptr, np.arange(n), side="right") - 1
    rows = np.repeat(np.arange(n), np.diff(sp.lo_ptr))
    assert np.all(stage_of[sp.lo_col] < stage_of[rows])
    rows = np.repeat(np.arange(n), np.diff(sp.up_ptr))
    assert np.all(stage_of[sp.up_col] > stage_of[rows])
    # blocks stay inside a stage
    assert np.all(stage_of[sp.row_block] == stage_of)
    assert np.all(stage_of[sp.row_block + sp.row_bsize - 1] == stage_of)
    assert sp.n_stages < 12


def test_cycle_codes():
    c = pa.HoneycombArray(3, 3)
    A = c.get_cycle_matrix()
    fp, fc, jc = cycle_codes(A, c.face_count(), c.junction_count())
    dense = np.zeros(A.shape)
    for f in range(A.shape[0]):
        for e in range(fp[f], fp[f + 1]):
            dense[f, fc[e] >> 1] = -1 if fc[e] & 1 else 1
        assert np.all(np.diff(fc[fp[f]:fp[f + 1]] >> 1) > 0)   # ascending junctions
    assert np.array_equal(dense, A.toarray())
    dense2 = np.zeros(A.shape)
    for j in range(A.shape[1]):
        for k in range(2):
            code = jc[2 * j + k]
            if code >= 0:
                dense2[code >> 1, j] = -1 if code & 1 else 1
    assert np.array_equal(dense2, A.toarray())


@pytest.mark.parametrize("BS", [64, 256, 1024])
@pytest.mark.parametrize("name", ["sq10", "hc7", "tri8", "sqL", "het", "squid"])
def test_sliced_layout_emulation(name, BS):
    """The exact face phase and staged solve the kernel runs, executed on the host
    from the sliced layouts, equals a direct solve of K J = A u."""
    c = _circuits()[name]
    plan = _plan(c, 0.05, leaf_size=8, block_size=BS)
    u = np.random.default_rng(2).standard_normal(c.junction_count())
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(plan.K), c.get_cycle_matrix() @ u)
    got = plan.emulate_face_and_solve(u)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * abs(want).max())


def test_sliced_layout_holed_array():
    full = pa.SquareArray(9, 9)
    x, y = full.get_node_coordinates()
    c = full.remove_nodes((x > 2) & (x < 6) & (y > 2) & (y < 6))
    plan = _plan(c, 0.05, block_size=64)
    u = np.random.default_rng(3).standard_normal(c.junction_count())
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(plan.K), c.get_cycle_matrix() @ u)
    np.testing.assert_allclose(plan.emulate_face_and_solve(u), want, rtol=1e-12,
                               atol=1e-12 * abs(want).max())


@pytest.mark.parametrize("store", ["lds", "global"])
@pytest.mark.parametrize("BS", [64, 1024])
@pytest.mark.parametrize("name", ["sq10", "hc7", "tri8", "sqL", "het", "squid"])
def test_plan_descriptors_pass_native_validation(name, BS, store):
    """Every plan the host builds passes the library's own index and offset checks
    (the ones jja_plan_create runs before any upload); no device is touched."""
    import ctypes as C
    from pyjjasim_amd import _native
    plan = _plan(_circuits()[name], 0.05, leaf_size=8, block_size=BS, face_store=store)
    assert plan.index_bytes == (4 if store == "global" else 2)
    d = plan.desc()
    NW = BS // 64
    assert all((plan.wave_ptr[w + 1] - plan.wave_ptr[w] - 4) % 3 == 0 for w in range(NW))
    assert _native.lib().jja_plan_validate(1, C.byref(d)) == 0, _native.lib().jja_last_error()


def test_native_validation_rejects_bad_programs():
    """A corrupted LDS offset, a wave program of odd length or waves that disagree
    on their barriers are refused (EINVAL) before anything runs."""
    import ctypes as C
    from pyjjasim_amd import _native
    lib = _native.lib()

    def fresh():
        return _plan(_circuits()["sq10"], 0.05, leaf_size=8, block_size=128)
    plan = fresh()
    k = next(i for i in range(len(plan.solve_slices)) if plan.solve_slices[i][4] > 0
             and not plan.solve_slices[i][5] & _native.SLICE_BLOCK)
    plan.ent_off[plan.solve_slices[k][1] + 3] = plan.Nf + plan.scratch_rows
    assert lib.jja_plan_validate(1, C.byref(plan.desc())) == _native.JJA_EINVAL
    assert b"offset" in lib.jja_last_error()
    # a global store's block-phase lane word (base | len << 28) past the store
    plan = _plan(_circuits()["sq10"], 0.05, leaf_size=8, block_size=128, face_store="global")
    k = next(i for i in range(len(plan.solve_slices)) if plan.solve_slices[i][5] & _native.SLICE_BLOCK)
    plan.ent_off[plan.solve_slices[k][1]] = (plan.Nf + plan.scratch_rows) | (1 << 28)
    assert lib.jja_plan_validate(1, C.byref(plan.desc())) == _native.JJA_EINVAL
    assert b"offset" in lib.jja_last_error()
    plan = fresh()
    plan._wave_ptr[1] -= 1
    assert lib.jja_plan_validate(1, C.byref(plan.desc())) == _native.JJA_EINVAL
    plan = fresh()
    w1 = [i for i in range(int(plan.wave_ptr[1]), int(plan.wave_ptr[2]) - 4)
          if plan._solve_rec[i].flags & _native.SLICE_BARRIER]
    plan._solve_rec[w1[0]].flags &= ~_native.SLICE_BARRIER
    assert lib.jja_plan_validate(1, C.byref(plan.desc())) == _native.JJA_EINVAL
    assert b"barrier" in lib.jja_last_error()


def test_large_circuit_plans_a_global_face_store():
    """A face vector beyond one workgroup's LDS is planned in global memory with
    32-bit indices, and the host emulation of that program still solves K J = A u."""
    c = pa.SquareArray(130, 130)
    plan = _plan(c, 0.05, block_size=1024)
    assert plan.global_store and plan.index_bytes == 4 and plan.ent_off.dtype == np.uint32
    with pytest.raises(ValueError, match="LDS"):
        _plan(c, 0.05, block_size=1024, face_store="lds")
    small = pa.SquareArray(8, 8)
    p2 = _plan(small, 0.05, block_size=256, face_store="global")
    u = np.random.default_rng(4).standard_normal(small.junction_count())
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(p2.K), small.get_cycle_matrix() @ u)
    np.testing.assert_allclose(p2.emulate_face_and_solve(u), want, rtol=1e-12, atol=1e-12 * abs(want).max())


@pytest.mark.parametrize("bs", [128, 256, 512, 1024])
def test_relayout_plan_validates_and_solves(bs):
    """A plan re-laid out for another block size keeps the factor and passes
    the library's checks; its wave programs still solve K J = A u."""
    import ctypes as C
    from pyjjasim_amd import _native
    a = pa.TriangularArray(9, 7)
    a.set_inductance(0.7)
    plan = _plan(a, 0.05, block_size=1024)
    p2 = plan.relayout(bs)
    assert p2.block_size == bs
    assert _native.lib().jja_plan_validate(1, C.byref(p2.desc())) == 0, _native.lib().jja_last_error()
    assert p2.solve_plan is plan.solve_plan and plan.block_size == 1024
    u = np.random.default_rng(bs).standard_normal(a.junction_count())
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(plan.K), a.get_cycle_matrix() @ u)
    np.testing.assert_allclose(p2.emulate_face_and_solve(u), want, rtol=1e-12,
                               atol=1e-12 * abs(want).max())


def test_plans_pickle_and_build_in_worker_processes():
    """Plans built in worker processes equal those built here, record arrays
    included (they travel by pickle and are rebuilt on arrival)."""
    import pickle
    import ctypes as C
    from pyjjasim_amd import _native
    from pyjjasim_amd.batch import build_plans
    circuits = [pa.SquareArray(12, 9), pa.HoneycombArray(6, 5), pa.TriangularArray(7, 7)]
    # current_base caches a SuperLU factor on the circuit: it must not travel
    probs = [pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=10,
                                     current_sources=c.current_base(0.0)[:, None, None]) for c in circuits]
    here = build_plans(probs, 256, workers=1)
    there = build_plans(probs, 256, workers=2)
    for a, b in zip(here, there):
        for name in ("ent_val", "ent_off", "slice_rows", "face_code_sliced", "junc_pos_code", "coef"):
            np.testing.assert_array_equal(getattr(a, name), getattr(b, name))
        assert a.solve_slices == b.solve_slices and a.face_slices == b.face_slices
        assert bytes(a._solve_rec) == bytes(b._solve_rec)
        assert _native.lib().jja_plan_validate(1, C.byref(b.desc())) == 0
        c2 = pickle.loads(pickle.dumps(a))
        assert bytes(c2._face_rec) == bytes(a._face_rec)


@pytest.mark.parametrize("bs,parts", [(64, 2), (256, 4), (1024, 2)])
def test_split_column_plans_validate_and_solve(bs, parts):
    """A column split over `parts` workgroups is laid out for parts * bs threads
    with the global face store; the library accepts it and its programs solve
    K J = A u."""
    import ctypes as C
    from pyjjasim_amd import _native
    c = _circuits()["tri8"]
    plan = _plan(c, 0.05, leaf_size=8, block_size=bs, parts=parts)
    assert plan.global_store and plan.column_threads == bs * parts
    d = plan.desc()
    assert d.block_size == bs * parts and d.parts == parts
    assert _native.lib().jja_plan_validate(1, C.byref(d)) == 0, _native.lib().jja_last_error()
    u = np.random.default_rng(5).standard_normal(c.junction_count())
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(plan.K), c.get_cycle_matrix() @ u)
    np.testing.assert_allclose(plan.emulate_face_and_solve(u), want, rtol=1e-12, atol=1e-12 * abs(want).max())
    p1 = _plan(c, 0.05, leaf_size=8, block_size=bs).relayout(bs, parts)
    assert p1.column_threads == bs * parts and p1.global_store


def test_split_column_plans_rejected_where_unsupported():
    import ctypes as C
    from pyjjasim_amd import _native
    c = _circuits()["sq10"]
    with pytest.raises(ValueError, match="global face store"):
        _plan(c, 0.05, block_size=256, parts=2, face_store="lds")
    with pytest.raises(ValueError, match="power of two"):
        _plan(c, 0.05, block_size=256, parts=3)
    plan = _plan(c, 0.05, block_size=256, parts=2)
    d = plan.desc()
    d.index_bytes = 2                    # split plans need the global store
    assert _native.lib().jja_plan_validate(1, C.byref(d)) == _native.JJA_EINVAL
    d = plan.desc()
    d.parts = 3
    assert _native.lib().jja_plan_validate(1, C.byref(d)) == _native.JJA_EINVAL


def test_balanced_parts_allocation():
    """The slowest problem gets its parts doubled while the device holds every
    workgroup; small problems stay at one part."""
    from types import SimpleNamespace
    from pyjjasim_amd.batch import balanced_parts
    # split cost 2 * (off-block + block-inverse entries) of the factor
    hosts = [SimpleNamespace(solve_plan=SimpleNamespace(lo_val=np.zeros(n // 2), inv_cm=np.zeros(0)), Nj=0,
                             uses_lattice=False)
             for n in (64000, 16000, 1000, 1000)]
    parts = balanced_parts(hosts, [1, 1, 1, 1], n_cu=16, max_parts=16)
    # 1,1,1,1 (4 workgroups) -> first 2, 4, 8 for the largest (cost 64000 / parts),
    # then 2 for the second; doubling the largest again would need 20 > 16
    assert parts == [8, 2, 1, 1]
    # with no room for splits nothing changes
    assert balanced_parts(hosts, [4, 4, 4, 4], n_cu=16) == [1, 1, 1, 1]


def test_mixed_part_plans_validate():
    """Problems of one plan may have different part counts (same workgroup
    size); once one is split, all need the global face store."""
    import ctypes as C
    from pyjjasim_amd import _native
    a = _plan(_circuits()["sq10"], 0.05, block_size=256, parts=4)
    b = _plan(_circuits()["hc7"], 0.05, block_size=256, face_store="global")
    descs = (_native.ProblemDesc * 2)(a.desc(), b.desc())
    assert _native.lib().jja_plan_validate(2, descs) == 0, _native.lib().jja_last_error()
    c = _plan(_circuits()["hc7"], 0.05, block_size=256)          # LDS store
    assert not c.global_store
    descs = (_native.ProblemDesc * 2)(a.desc(), c.desc())
    assert _native.lib().jja_plan_validate(2, descs) == _native.JJA_EINVAL
    d = _plan(_circuits()["hc7"], 0.05, block_size=128, face_store="global")
    descs = (_native.ProblemDesc * 2)(a.desc(), d.desc())
    assert _native.lib().jja_plan_validate(2, descs) == _native.JJA_EINVAL


@pytest.mark.parametrize("store", ["lds", "global"])
@pytest.mark.parametrize("BS,parts", [(64, 1), (256, 1), (1024, 1), (256, 2)])
@pytest.mark.parametrize("name", ["sq10", "hc7", "tri8", "sqL", "het", "squid"])
def test_wave_local_stages_emulation(name, BS, parts, store):
    """Lowest stages solved wave-locally (whole subtrees per wave, no barrier
    between those stages): the plans validate, and the host emulation, which
    runs each wave's records in order and fails on any element one wave writes
    and another touches between the same barriers, solves K J = A u."""
    import ctypes as C
    from pyjjasim_amd import _native
    if parts > 1 and store == "lds":
        pytest.skip("split columns use the global face store")
    c = _circuits()[name]
    u = np.random.default_rng(6).standard_normal(c.junction_count())
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(_plan(c, 0.05, leaf_size=8).K),
                                       c.get_cycle_matrix() @ u)
    base = _plan(c, 0.05, leaf_size=8, block_size=BS, parts=parts, face_store=store, local_stages=0)
    for k in (1, 2, 99):
        plan = base.relayout(BS, local_stages=k)
        assert plan.local_stages == min(k, plan.solve_plan.n_stages)
        assert plan.scratch_rows == max(plan.solve_plan.max_stage_rows,
                                        int(plan.solve_plan.stage_ptr[max(plan.local_stages,
                                                                          plan.part_stages)]))
        assert _native.lib().jja_plan_validate(1, C.byref(plan.desc())) == 0, _native.lib().jja_last_error()
        np.testing.assert_allclose(plan.emulate_face_and_solve(u), want, rtol=1e-12,
                                   atol=1e-12 * abs(want).max())


def test_wave_local_barrier_count():
    """k wave-local stages replace their 4k - 1 phase barriers (the leaves'
    first phase is skipped) by one per pass."""
    c = pa.TriangularArray(14, 14)
    base = _plan(c, 0.05, leaf_size=8, block_size=256, local_stages=0)

    def barriers(plan):
        return sum(1 for r in plan.wave_program(0) if r[5] & 0x400)
    n0 = barriers(base)
    assert n0 == 4 * base.solve_plan.n_stages - 1
    for k in (1, 3):
        assert barriers(base.relayout(256, local_stages=k)) == n0 - (4 * k - 1) + 2


class _MisownedPlan(CircuitPlan):
    """Planted mistake: the leaves of one subtree go to another wave than their
    separators, so two waves share rows between barriers."""

    def _wave_owner(self, NW, n_low):
        owner = super()._wave_owner(NW, n_low)
        if owner is not None:
            leaves = int(self.solve_plan.stage_ptr[1])
            moved = owner[:leaves] == owner[0]
            owner[:leaves][moved] = (owner[0] + 1) % NW
        return owner


def test_wave_local_race_is_detected():
    c = _circuits()["tri8"]
    plan = _MisownedPlan(c, 0.05, leaf_size=8, block_size=256, local_stages=3)
    u = np.random.default_rng(7).standard_normal(c.junction_count())
    with pytest.raises(AssertionError, match="another wave"):
        plan.emulate_face_and_solve(u)


def test_local_stage_choice_and_lds_fit():
    """auto: none with one workgroup per column, some for split columns. A
    count whose scratch does not fit LDS moves the store to global memory, or
    is refused when LDS is forced."""
    c = _circuits()["tri8"]
    assert _plan(c, 0.05, leaf_size=8, block_size=256).local_stages == 0
    assert _plan(c, 0.05, leaf_size=8, block_size=256, parts=2).local_stages > 0
    with pytest.raises(ValueError, match="local_stages"):
        _plan(c, 0.05, local_stages=-1)
    big = pa.SquareArray(103, 103)          # 10404 faces: LDS holds the store only in step
    p0 = _plan(big, 0.05, block_size=1024)
    assert not p0.global_store and p0.local_stages == 0
    pall = p0.relayout(1024, local_stages=99)
    assert pall.global_store and pall.scratch_rows > p0.scratch_rows
    with pytest.raises(ValueError, match="LDS"):
        p0.relayout(1024, face_store="lds", local_stages=99)


def test_column_pair_plans_validate():
    """A split plan worked two columns at a time keeps its programs (half the
    workgroup size, twice the parts); the library accepts it, and refuses
    pairs without split columns, an unknown value or plans that mix them."""
    import ctypes as C
    from pyjjasim_amd import _native
    c = _circuits()["tri8"]
    split = _plan(c, 0.05, leaf_size=8, block_size=512, parts=2)
    pair = split.paired()
    assert (pair.block_size, pair.parts, pair.column_threads) == (256, 4, split.column_threads)
    assert pair.solve_slices is split.solve_slices
    d = pair.desc()
    assert d.column_pairs == 1
    assert _native.lib().jja_plan_validate(1, C.byref(d)) == 0, _native.lib().jja_last_error()
    d.column_pairs = 2
    assert _native.lib().jja_plan_validate(1, C.byref(d)) == _native.JJA_EINVAL
    one = _plan(c, 0.05, leaf_size=8, block_size=256, face_store="global").desc()
    one.column_pairs = 1                          # pairs need split columns
    assert _native.lib().jja_plan_validate(1, C.byref(one)) == _native.JJA_EINVAL
    mixed = (_native.ProblemDesc * 2)(pair.desc(), _plan(c, 0.05, leaf_size=8, block_size=256,
                                                                  parts=4).desc())
    assert _native.lib().jja_plan_validate(2, mixed) == _native.JJA_EINVAL
    with pytest.raises(ValueError, match="column pairs"):
        _plan(c, 0.05, leaf_size=8, block_size=256).paired()
    assert not split.relayout(512, 2).column_pairs


def _part_local_plan(cls=CircuitPlan):
    c = pa.SquareArray(64, 64)
    plan = cls(c, 0.05, block_size=512, parts=2, solver="cholesky")
    assert plan.part_stages > plan.local_stages > 0
    return c, plan


def test_part_local_stages_emulation_and_validation():
    """A split column whose middle stages are part-local (whole subtrees per
    workgroup, workgroup barriers between their phases): the plan validates,
    every wave meets the same barrier sequence, and the emulation (workgroup
    barriers join only a part's waves) solves K J = A u."""
    import ctypes as C
    from pyjjasim_amd import _native
    c, plan = _part_local_plan()
    assert _native.lib().jja_plan_validate(1, C.byref(plan.desc())) == 0, _native.lib().jja_last_error()
    kinds = [[r[5] & (_native.SLICE_BARRIER | _native.SLICE_WG_BARRIER) for r in plan.wave_program(w)
              if r[5] & (_native.SLICE_BARRIER | _native.SLICE_WG_BARRIER)]
             for w in range(plan.column_threads // 64)]
    assert all(k == kinds[0] for k in kinds) and _native.SLICE_WG_BARRIER in kinds[0]
    u = np.random.default_rng(8).standard_normal(c.junction_count())
    want = scipy.sparse.linalg.spsolve(scipy.sparse.csc_matrix(plan.K), c.get_cycle_matrix() @ u)
    np.testing.assert_allclose(plan.emulate_face_and_solve(u), want, rtol=1e-12, atol=1e-12 * abs(want).max())
    # every stage in step: no workgroup barriers, the same solution
    flat = plan.relayout(512, local_stages=0)
    assert flat.part_stages == 0
    np.testing.assert_allclose(flat.emulate_face_and_solve(u), want, rtol=1e-12, atol=1e-12 * abs(want).max())


class _MisownedPartPlan(CircuitPlan):
    """Planted mistake: the part-local rows of the lowest part-local stage that
    part 0 owns go to part 1, away from their wave-local descendants."""

    def _owners(self, NW, n_wave, n_part):
        wave_of, part_of = super()._owners(NW, n_wave, n_part)
        if part_of is not None:
            r0 = int(self.solve_plan.stage_ptr[n_wave])
            r1 = int(self.solve_plan.stage_ptr[n_wave + 1])
            part_of = part_of.copy()
            sel = np.arange(r0, r1)[part_of[r0:r1] == 0]
            part_of[sel] = 1
        return wave_of, part_of


def test_part_local_race_is_detected():
    c, plan = _part_local_plan(_MisownedPartPlan)
    u = np.random.default_rng(9).standard_normal(c.junction_count())
    with pytest.raises(AssertionError, match="another (part|wave)"):
        plan.emulate_face_and_solve(u)


def test_native_validation_rejects_barrier_sequences_that_differ():
    import ctypes as C
    from pyjjasim_amd import _native
    _, plan = _part_local_plan()
    w1 = [i for i in range(int(plan.wave_ptr[1]), int(plan.wave_ptr[2]) - 4)
          if plan._solve_rec[i].flags & _native.SLICE_WG_BARRIER]
    rec = plan._solve_rec[w1[0]]
    rec.flags = (rec.flags & ~_native.SLICE_WG_BARRIER) | _native.SLICE_BARRIER
    assert _native.lib().jja_plan_validate(1, C.byref(plan.desc())) == _native.JJA_EINVAL


@pytest.mark.parametrize("kind", ["cycle", "cut"])
@pytest.mark.parametrize("lattice", ["square", "honeycomb", "triangular"])
def test_matrix_plan_solve_program_matches_scipy(kind, lattice):
    """A plan made from a matrix alone (the cycle system A A^T or the grounded
    cut system Mr Mr^T): the kernel's solve program, run record by record on the
    host, gives scipy's solution to rounding, and the descriptor validates."""
    import ctypes as C
    import scipy.sparse.linalg
    from pyjjasim_amd import _native
    from pyjjasim_amd.solver_plan import MatrixPlan
    c = {"square": pa.SquareArray(13, 9), "honeycomb": pa.HoneycombArray(6, 5),
         "triangular": pa.TriangularArray(7, 6)}[lattice]
    if kind == "cycle":
        A = c.get_cycle_matrix()
        K = (A @ A.T).tocsc()
        xy = np.stack(c.get_face_centroids(), axis=1)
    else:
        Mr = c._Mr()
        K = (Mr @ Mr.T).tocsc()
        x, y = c.get_node_coordinates()
        xy = np.stack([x[:-1], y[:-1]], axis=1)
    mp = MatrixPlan(K, xy, leaf_size=8)
    b = np.random.default_rng(3).standard_normal(K.shape[0])
    want = scipy.sparse.linalg.spsolve(K.astype(float), b)
    got = mp.emulate_solve(b)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert _native.lib().jja_plan_validate(1, C.byref(mp.desc())) == 0, _native.lib().jja_last_error()
