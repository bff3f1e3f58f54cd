"""Annealing on the device: AnnealingProblem (time_evolution.py:1070-1191).

The reference runs ``interval_count`` separate ``compute()`` calls of
``interval_steps`` steps, downloads every step's vortex configuration to
measure the vortex mobility on the host, adjusts each column's temperature,
and starts the next interval from the last phases (time_evolution.py:1142-1174).

Here one device plan integrates every interval: each interval is one launch of
the step kernel whose phases go to a ring of ``interval_steps`` device slots;
the vortex kernel counts every slot's vortices and the mobility
sum_faces sum_steps |n_{s+1} - n_s| of each column is reduced on the device
(integer sums, exact), so per interval only W numbers reach the host, where the
reference's temperature rule (time_evolution.py:1128-1140) sets the next
temperatures. Each interval restarts from theta_-1 = theta_-2 = the previous
interval's last phases, as the reference does. The noise amplitude
sqrt(2 T_w Rv_j) of every column follows its temperature.

With ``noise="reference"`` each interval replays the reference's legacy numpy
draws of a fresh compute() (its step index restarting at 0), so a seeded run
follows the reference's temperature profile exactly; with ``noise="philox"``
(default) the in-kernel stream runs on the absolute step count, so every
interval draws new deviates. The closing part is the reference's: five runs
of ``interval_steps`` steps at T = 0 with half the time step, then a static
(Newton) solve per column for the final vortex configuration
(time_evolution.py:1176-1191).
"""
from __future__ import annotations

import numpy as np

from . import _native
from .circuit import Circuit
from .engine import Engine, reference_noise, _torch
from .time_evolution import TimeEvolutionProblem

__all__ = ["AnnealingProblem"]


class AnnealingProblem:
    """Anneal a circuit by lowering the temperature under vortex-mobility
    feedback (time_evolution.py:1070-1123). Parameters are the reference's;
    ``noise`` and ``seed`` select the noise stream as in TimeEvolutionProblem."""

    def __init__(self, circuit: Circuit, time_step=0.5, interval_steps=10, external_flux=0.0,
                 current_sources=0, problem_count=1, interval_count=1000, vortex_mobility=0.001,
                 start_T=1.0, T_factor=1.03, *, noise="philox", seed=None):
        self.circuit = circuit
        self.time_step = time_step
        self.interval_steps = interval_steps
        self.interval_count = interval_count
        self.vortex_mobility = vortex_mobility
        self.current_sources = current_sources
        self.external_flux = external_flux
        self.problem_count = problem_count
        self.T = start_T * np.ones((1, self.problem_count, 1))
        self.T_factor = T_factor
        if noise not in ("philox", "reference"):
            raise ValueError("noise must be 'philox' or 'reference'")
        self.noise = noise
        self.seed = seed

    def get_vortex_mobility(self, n):
        """Mobility of consecutive vortex configurations n (Nf, W, steps), host
        form (time_evolution.py:1123-1128); the annealing loop uses the device
        reduction with the same integer sums and the same division."""
        Nf = self.circuit.face_count()
        return np.sum(np.sum(np.abs(np.diff(n, axis=2)), axis=2), axis=0) / \
            (Nf * self.time_step * (self.interval_count - 1))

    def _temperature_adjustment(self, vortex_mobility, iteration):
        v = self.vortex_mobility
        upper = v[iteration] if (np.array(v)).size == self.interval_count else \
            v * ((self.interval_count - iteration) / self.interval_count) ** 1.5
        factor = (vortex_mobility > upper) * (1 / self.T_factor) + (vortex_mobility <= upper) * self.T_factor
        self.T *= factor[..., None]

    # ------------------------------------------------------------------ run
    def _engine(self, dt, n_intervals, T0):
        """A plan for n_intervals intervals of interval_steps steps at time step
        dt; theta goes to a ring of interval_steps device slots."""
        torch = _torch()
        S = self.interval_steps
        Nt = S * n_intervals
        f = np.atleast_1d(self.external_flux)[:, None, None]
        prob = TimeEvolutionProblem(self.circuit, time_step=dt, time_step_count=Nt, external_flux=f,
                                    current_sources=self.current_sources, temperature=T0,
                                    store_time_steps=[Nt - 1], store_current=False, store_voltage=False,
                                    noise=self.noise, seed=self.seed, steps_per_launch=S,
                                    max_resident_output_bytes=1 << 62)
        W = prob.problem_count
        if W != self.problem_count:
            raise ValueError("current_sources and external_flux must match problem_count columns")
        eng = Engine(prob, np.zeros(Nt, dtype=bool), np.zeros(Nt, dtype=bool))
        ring = torch.empty((S, W, eng.Nj), dtype=torch.float64, device=eng.dev)
        slots = torch.from_numpy(np.tile(np.arange(S, dtype=np.int32), n_intervals)).to(eng.dev)
        eng.run.th_slot, eng.run.slot_t0, eng.run.th_out = slots.data_ptr(), 0, ring.data_ptr()
        eng._ring, eng._slots = ring, slots
        return eng

    def _set_temperature(self, eng, T_w):
        """Noise amplitude sqrt(2 T_w Rv_j) of every junction and column (the
        expression of time_evolution.py:538), as an (Nj, W) device source."""
        torch = eng.torch
        sig = np.sqrt(2.0 * np.asarray(T_w, dtype=np.float64)[None, :] * eng.host.Rv[:, None])
        eng._sig = torch.from_numpy(np.ascontiguousarray(sig.T)).to(eng.dev)      # [w][j]
        src = _native.Source()
        src.base = eng._sig.data_ptr()
        src.s_i, src.s_w, src.s_t, src.t0, src.zero = 1, eng.Nj, 0, 0, 0
        eng.run.sigma = src

    def _interval(self, eng, th):
        """Start an interval from theta_-1 = theta_-2 = th (W, Nj) and run it."""
        eng.theta[0].copy_(th)
        eng.theta[1].copy_(th)
        if eng.ref_noise:
            eng.draw = reference_noise(eng.Nj, eng.W)
            t0 = eng.t
            base = eng.draw
            eng.draw = lambda t: base(t - t0)          # the draw index restarts at 0
        eng.launch(self.interval_steps)
        return eng._ring[-1].clone()

    def _mobility(self, eng):
        """Vortex mobility of the ring's steps, per column, reduced on the device."""
        torch = eng.torch
        S, W, Nf = self.interval_steps, eng.W, eng.Nf
        n = torch.empty((S, W, Nf), dtype=torch.int32, device=eng.dev)
        stream = torch.cuda.current_stream().cuda_stream
        _native.check(eng.lib.jja_vortex(eng.handle, eng.index, eng._ring.data_ptr(), S * W, n.data_ptr(),
                                         stream))
        moves = (n[1:] - n[:-1]).abs().to(torch.int64).sum(dim=(0, 2)).cpu().numpy()
        return moves / (Nf * self.time_step * (self.interval_count - 1))

    def compute(self):
        """Returns (status (W,), configurations (W StaticConfiguration),
        temperature_profiles (interval_count, W)) (time_evolution.py:1142-1191)."""
        torch = _torch()
        Nj, W = self.circuit.junction_count(), self.problem_count
        temperature_profiles = np.zeros((self.interval_count, W))
        eng = self._engine(self.time_step, self.interval_count, self.T)
        try:
            th = torch.zeros((W, Nj), dtype=torch.float64, device=eng.dev)
            for i in range(self.interval_count):
                self._set_temperature(eng, self.T[0, :, 0])
                th = self._interval(eng, th)
                self._temperature_adjustment(self._mobility(eng), i)
                temperature_profiles[i, :] = self.T[0, :, 0]
            _native.check(eng.lib.jja_plan_status(eng.handle))
        finally:
            eng.close()
        # five runs at T = 0 with half the time step
        end = self._engine(self.time_step / 2, 5, np.zeros((1, W, 1)))
        try:
            for i in range(5):
                th = self._interval(end, th)
            n = torch.empty((W, end.Nf), dtype=torch.int32, device=end.dev)
            _native.check(end.lib.jja_vortex(end.handle, end.index, th.data_ptr(), W, n.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream))
            n = n.cpu().numpy().T.astype(int)
        finally:
            end.close()
        f = np.atleast_1d(self.external_flux)[:, None, None]
        prob = TimeEvolutionProblem(self.circuit, time_step_count=self.interval_steps,
                                    time_step=self.time_step / 2, external_flux=f,
                                    current_sources=self.current_sources, store_current=False,
                                    store_voltage=False)
        data = [prob.get_static_problem(n[:, p], problem_nr=0, time_step=0).compute() for p in range(W)]
        configurations = [d[0] for d in data]
        status = np.array([d[1] for d in data])
        self.final_theta = th.cpu().numpy().T
        return status, configurations, temperature_profiles
