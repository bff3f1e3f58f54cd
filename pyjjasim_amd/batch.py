"""Many time-evolution problems in one persistent launch per chunk of steps.

``TimeEvolutionBatch(problems).compute()`` returns one TimeEvolutionResult per
problem, each identical to ``problem.compute()``: the problems (any circuits,
sizes and column counts; one stencil width and step count) share one libjja
plan, and every launch advances the columns of all of them, one workgroup per
column. Circuits whose face vector fits in LDS keep it there; larger ones use
the global face store, in a second launch of the same step range.

``autotune=True`` picks the workgroup size and the launch length by measuring:
each candidate workgroup size is timed on a short run of the batch (the factors
are computed once; only the record layouts change), and the launch length is
the shortest whose measured launch overhead stays under 1% of its run time.
"""
from __future__ import annotations

import time

import numpy as np

from .engine import BatchEngine, _torch
from .solver_plan import CircuitPlan
from .time_evolution import TimeEvolutionProblem, _finish, _masks

__all__ = ["TimeEvolutionBatch"]


class TimeEvolutionBatch:
    def __init__(self, problems, block_size=None, steps_per_launch=None, autotune=False,
                 block_sizes=(256, 512, 1024), tune_steps=40):
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
        self.tuning = None

    # --------------------------------------------------------------- tuning
    def _engine(self, hosts, masks):
        return BatchEngine(self.problems, masks, hosts=hosts)

    def _tune(self, masks):
        """Time each candidate workgroup size on tune_steps steps; returns
        (block_size, hosts, record)."""
        torch = _torch()
        base = [CircuitPlan(p.circuit, p._dt(), p.stencil_width, leaf_size=p.leaf_size,
                            block_size=self.block_sizes[0],
                            face_store=getattr(p, "face_store", "auto")) for p in self.problems]
        steps = min(self.tune_steps, self.problems[0]._Nt())
        record = {}
        best = None
        for bs in self.block_sizes:
            hosts = [h if h.block_size == bs else h.relayout(bs) for h in base]
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
            record[bs] = dt
            if best is None or dt < record[best[0]]:
                best = (bs, hosts)
        return best[0], best[1], record

    def _launch_length(self, eng):
        """Shortest launch length whose measured fixed cost per launch stays below
        1% of the launch (at least 100 steps, at most 4000)."""
        torch = _torch()
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
    def compute(self):
        masks = [_masks(p) for p in self.problems]
        core_masks = [(m[2], m[3]) for m in masks]
        hosts = None
        if self.autotune:
            bs, hosts, record = self._tune(core_masks)
            self.block_size = bs
            self.tuning = {"block_size_s_per_step": record, "block_size": bs}
        eng = BatchEngine(self.problems, core_masks, block_size=self.block_size, hosts=hosts)
        try:
            if self.autotune and not self.steps_per_launch:
                spl = self._launch_length(eng)
                eng.close()
                eng = BatchEngine(self.problems, core_masks, block_size=self.block_size, hosts=hosts)
                eng.steps_per_launch = spl
                self.tuning["steps_per_launch"] = spl
            elif self.steps_per_launch:
                eng.steps_per_launch = int(self.steps_per_launch)
            eng.advance(eng.Nt)
            cores = eng.results()
        finally:
            eng.close()
        return [_finish(p, m, *c) for p, m, c in zip(self.problems, masks, cores)]
