"""Many time-evolution problems in one launch per chunk of steps.

``TimeEvolutionBatch(problems).compute()`` returns one TimeEvolutionResult per
problem, each identical to ``problem.compute()`` run with the layout the batch
gave it (``layouts``: workgroup size, parts per column, face store; the factor
gives the same bits in any layout, while the lattice solve's rounding depends
on the face store and the part count): the problems (any circuits,
sizes and column counts; one stencil width and step count) share one libjja
plan, and every launch advances the columns of all of them. By default a
column is one workgroup: circuits whose face vector fits in LDS keep it there,
larger ones use the global face store, in the same launch. With parts > 1 (or
"balanced", a part count per problem) large problems' columns are split over
several workgroups, and every face store is global. Workgroups are ordered
largest problem first, so the hardware's in-order dispatch is a longest-first
list schedule.

Random inputs: with Philox noise each problem uses its own ``seed`` (a problem
without one draws it from numpy's global RNG when the batch's engine is made,
problem by problem). With ``noise="reference"`` the launches draw every
problem's normals from numpy's global RNG, problem by problem within each chunk
of steps, so a batch of one problem repeats its compute() exactly after the same
np.random.seed, and a larger batch uses the same stream in that interleaved
order.

``autotune=True`` picks the workgroup size, the parts per column and the launch
length by measuring: each candidate (workgroup size, parts) is timed on a short
run of the batch (the factors are computed once; only the record layouts
change), and the launch length is the shortest whose measured launch overhead
stays under 1% of its run time. With parts > 1 every column is worked by that
many workgroups meeting at column barriers (all face stores global).
"""
from __future__ import annotations

import heapq
import multiprocessing
import os
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from .engine import BatchEngine, _torch, default_block_size, device_compute_units
from .solver_plan import CircuitPlan
from .time_evolution import TimeEvolutionProblem, _finish, _masks

__all__ = ["TimeEvolutionBatch", "build_plans"]


_THREAD_VARS = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "BLIS_NUM_THREADS")


def _split_cost(h):
    """Work of one column step of a problem in a split launch, in factor-entry
    units: the factor entries both passes stream, or for a lattice plan its
    transforms' multiply-adds (on the matrix cores, about twelve per entry's
    time) and passes over the faces; plus the junction work. (SquareArray(512,
    512): lattice 11.9M, factor 27.4M units; measured 0.336 and 0.966 ms per
    step.)"""
    if h.uses_lattice:
        lat = h.lattice
        return float(lat.R * lat.C * lat.ES / 12.0 + 8 * h.Nf + 8 * h.Nj)
    sp = h.solve_plan
    return float(2 * (len(sp.lo_val) + len(sp.inv_cm)) + 8 * h.Nj)


def balanced_parts(hosts, columns, n_cu, max_parts=16):
    """Parts per problem for one split launch: all start at one, and the problem
    whose columns take longest (split cost / parts) has its parts doubled
    while every workgroup of the launch still fits on the device."""
    parts = [1] * len(hosts)
    used = sum(columns)
    heap = [(-_split_cost(h), i) for i, h in enumerate(hosts)]
    heapq.heapify(heap)
    while heap:
        _, i = heapq.heappop(heap)
        if parts[i] >= max_parts or used + columns[i] * parts[i] > n_cu:
            break                      # the slowest problem cannot speed up further
        used += columns[i] * parts[i]
        parts[i] *= 2
        heapq.heappush(heap, (-_split_cost(hosts[i]) / parts[i], i))
    return parts


def _plan_one(args):
    circuit, dt, sw, leaf, bs, store, local, solver = args
    return CircuitPlan(circuit, dt, sw, leaf_size=leaf, block_size=bs, face_store=store,
                       local_stages=local, solver=solver)


def build_plans(problems, block_size, workers=None):
    """Host plans (ordering, factor, record layouts) of the problems' circuits,
    built in up to ``workers`` processes (default: the CPUs this process may run
    on, at most 15, so that with this process at most 16 processes hold the
    device runtime; one process when the batch is small). Worker processes
    only run host code; they are started fresh (spawn), never forked from a
    process that may hold a device context."""
    jobs = [(p.circuit, p._dt(), p.stencil_width, p.leaf_size, block_size,
             getattr(p, "face_store", "auto"), getattr(p, "local_stages", "auto"),
             getattr(p, "solver", "auto")) for p in problems]
    size = sum(p.circuit._Nj() for p in problems)
    if workers is None:
        workers = min(15, len(os.sched_getaffinity(0)), len(jobs)) if size > 200_000 else 1
    if workers <= 1:
        return [_plan_one(j) for j in jobs]
    # largest first, so the pool's tail is short
    order = sorted(range(len(jobs)), key=lambda i: -jobs[i][0]._Nj())
    ctx = multiprocessing.get_context("spawn")
    # one thread per worker: the workers' BLAS and OpenMP pools would otherwise
    # each take every CPU (workers read this environment when they start)
    saved = {k: os.environ.get(k) for k in _THREAD_VARS}
    os.environ.update({k: "1" for k in _THREAD_VARS})
    try:
        with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
            built = list(ex.map(_plan_one, [jobs[i] for i in order]))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    out = [None] * len(jobs)
    for i, plan in zip(order, built):
        out[i] = plan
    return out


class TimeEvolutionBatch:
    """problems: TimeEvolutionProblem objects (one step count and stencil width).
    block_size: threads per workgroup (default sized by the largest circuit).
    steps_per_launch: steps per launch (default the problems' own).
    autotune: time the (workgroup size, parts) candidates and the launch length.
    parts: workgroups per column, 1, a power of two, or "balanced" (per problem).
    workers: processes that build the host plans (default by batch size)."""

    def __init__(self, problems, block_size=None, steps_per_launch=None, autotune=False,
                 block_sizes=(256, 512, 1024), tune_steps=40, workers=None, parts=1, max_parts=16):
        self.problems = list(problems)
        if not self.problems:
            raise ValueError("TimeEvolutionBatch needs at least one problem")
        if not all(isinstance(p, TimeEvolutionProblem) for p in self.problems):
            raise TypeError("TimeEvolutionBatch takes TimeEvolutionProblem objects")
        self.block_size = block_size
        self.steps_per_launch = steps_per_launch
        self.autotune = autotune
        self.block_sizes = tuple(block_sizes)
        self.tune_steps = int(tune_steps)
        self.workers = workers
        self.parts = parts if parts == "balanced" else int(parts)
        self.max_parts = int(max_parts)
        self.tuning = None
        self.solvers = None
        self.layouts = None

    # --------------------------------------------------------------- tuning
    def _engine(self, hosts, masks):
        return BatchEngine(self.problems, masks, hosts=hosts)

    def _candidates(self):
        """(workgroup size, parts) pairs to time: every block size with one
        workgroup per column, and 1024-thread workgroups with 2, 4, ... parts per
        column (all face stores global) while every part stays resident."""
        out = [(bs, 1) for bs in self.block_sizes]
        if any(p.stencil_width != 3 or getattr(p, "face_store", "auto") == "lds" or
               getattr(p, "solver", "auto") == "lattice" for p in self.problems):
            return out
        cols = sum(p.problem_count for p in self.problems)
        n_cu = device_compute_units()
        parts = 2
        while parts <= self.max_parts and parts * cols <= n_cu:
            out.append((1024, parts))
            parts *= 2
        if len(self.problems) > 1:
            out.append((1024, "balanced"))
        return out

    def _hosts_for(self, base, bs, parts):
        """Plans laid out for workgroups of bs threads and the given parts: an
        int for every problem, or "balanced" (balanced_parts). With any split
        column, every problem uses the global face store (lattice plans keep the
        lattice solve)."""
        if parts == "balanced":
            cols = [p.problem_count for p in self.problems]
            per = balanced_parts(base, cols, device_compute_units(), self.max_parts)
        else:
            per = [int(parts)] * len(base)
        split = max(per) > 1
        if split:
            return [h if (h.block_size, h.parts) == (bs, q) and h.global_store
                    else h.relayout(bs, q, "global") for h, q in zip(base, per)]
        return [h if (h.block_size, h.parts) == (bs, q) else h.relayout(bs, q) for h, q in zip(base, per)]

    def _tune(self, masks):
        """Time each candidate (workgroup size, parts) on tune_steps steps;
        returns (block_size, parts, hosts, record)."""
        torch = _torch()
        base = build_plans(self.problems, self.block_sizes[0], self.workers)
        # a warm-up step and the timed steps, within the problems' horizon
        steps = min(self.tune_steps, self.problems[0]._Nt() - 1)
        if steps < 1:
            return self.block_sizes[0], 1, base, {}
        record = {}
        best = None
        for bs, parts in self._candidates():
            hosts = self._hosts_for(base, bs, parts)
            eng = self._engine(hosts, masks)
            try:
                eng.launch(1)                            # first launch warms up
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                eng.launch(steps)
                torch.cuda.synchronize()
                dt = (time.perf_counter() - t0) / steps
            finally:
                eng.close()
            record[(bs, parts)] = dt
            if best is None or dt < record[best[:2]]:
                best = (bs, parts, hosts)
        return best[0], best[1], best[2], record

    def _launch_length(self, eng):
        """Shortest launch length whose measured fixed cost per launch stays below
        1% of the launch (at least 100 steps, at most 4000). The measurement takes
        24 steps; a shorter horizon runs in one launch."""
        if eng.Nt < 24:
            return max(1, eng.Nt)
        torch = _torch()
        eng.launch(1)                 # the engine's first launch also sets it up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            eng.launch(1)
        torch.cuda.synchronize()
        per1 = (time.perf_counter() - t0) / 3
        t0 = time.perf_counter()
        eng.launch(20)
        torch.cuda.synchronize()
        per20 = time.perf_counter() - t0
        step = max((per20 - per1) / 19, 1e-9)
        fixed = max(per1 - step, 0.0)
        return int(min(4000, max(100, 100 * fixed / step)))

    # ---------------------------------------------------------------- run
    def make_engine(self, masks=None):
        """The batch's BatchEngine at step 0, tuned when autotune is set (the
        caller advances and closes it). masks: per problem (theta, current)
        store masks over the time steps; by default each problem's own."""
        if masks is None:
            masks = [_masks(p)[2:4] for p in self.problems]
        # tuning launches draw from numpy's global RNG (Philox seeds, reference
        # noise); the run itself starts from the state the caller left, so its
        # results equal each problem's own compute() after the same seeding
        rng_state = np.random.get_state()
        if self.autotune:
            bs, parts, hosts, record = self._tune(masks)
            self.block_size, self.parts = bs, parts
            self.tuning = {"s_per_step": {f"{b}x{q}": t for (b, q), t in record.items()},
                           "block_size": bs, "parts": parts,
                           "parts_per_problem": [h.parts for h in hosts]}
        else:
            self.block_size = self.block_size or max(
                default_block_size(p.circuit._Nj(), p.circuit._Nf()) for p in self.problems)
            hosts = build_plans(self.problems, self.block_size, self.workers)
            if self.parts == "balanced" or self.parts > 1:
                hosts = self._hosts_for(hosts, self.block_size, self.parts)
        np.random.set_state(rng_state)
        # how each problem's face system is solved in this batch, and its
        # layout (workgroup size, parts per column, face store)
        self.solvers = ["lattice" if h.uses_lattice else "cholesky" for h in hosts]
        self.layouts = [(h.block_size, h.parts, "global" if h.global_store else "lds") for h in hosts]
        eng = BatchEngine(self.problems, masks, block_size=self.block_size, hosts=hosts)
        try:
            if self.autotune and not self.steps_per_launch:
                spl = self._launch_length(eng)
                eng.close()
                np.random.set_state(rng_state)
                eng = BatchEngine(self.problems, masks, block_size=self.block_size, hosts=eng.hosts)
                eng.steps_per_launch = spl
                self.tuning["steps_per_launch"] = spl
            elif self.steps_per_launch:
                eng.steps_per_launch = int(self.steps_per_launch)
        except Exception:
            eng.close()
            raise
        return eng

    def compute(self):
        masks = [_masks(p) for p in self.problems]
        eng = self.make_engine([(m[2], m[3]) for m in masks])
        try:
            eng.advance(eng.Nt)
            cores = eng.results()
        finally:
            eng.close()
        return [_finish(p, m, *c) for p, m, c in zip(self.problems, masks, cores)]
