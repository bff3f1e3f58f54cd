"""Drop-in TimeEvolutionProblem / TimeEvolutionResult running on the GPU.

Same constructor, input semantics, storage rules, results and exceptions as the
reference (time_evolution.py:20-1067). ``compute()`` keeps the reference's
pre/post-processing in Python (``time_evolution``, time_evolution.py:422-458) and
replaces its core loop (``time_evolution_core``, :461-582) by libjja's persistent
HIP kernel: one workgroup per problem column integrates a chunk of steps, the face
system is solved with the staged Cholesky schedule (or, for a uniform rectangular
lattice, by sine transforms and tridiagonal solves), and vortex configurations of
stored steps are counted on the device.

Extensions (keyword-only, defaults keep reference behaviour):
  noise="philox" | "reference"
      "philox": seeded counter-based Gaussian stream computed in the kernel
      (reproducible for a given ``seed``; statistically equal to the reference's
      draws, eta = sqrt(2 T Rv) Z);
      "reference": the host replays the reference's legacy numpy RNG call sequence
      (np.random.randn / permutation, time_evolution.py:533-537) and streams the
      draws to the device, giving the reference's exact trajectories at T > 0.
  seed=None        Philox key; None draws one from numpy's global RNG.
  steps_per_launch Steps integrated per kernel launch (default: whole run in
                   launches of at most 2000 steps).
  block_size       Threads per workgroup (64..1024; default chosen by size).
  leaf_size        Largest leaf block of the nested-dissection factor schedule.
  face_store       "auto" (default): each column keeps its face vector in LDS when
                   it fits (up to about 10^4 faces) and in global memory otherwise;
                   "lds" or "global" force one.
  parts            Workgroups per column for circuits with a global face store:
                   None (default) picks the largest power of two up to 16 that
                   keeps all columns' workgroups resident on the device; 1 keeps
                   one workgroup per column. With the factor, results do not
                   depend on it; the lattice solve's rounding does (1e-12).
  local_stages     Lowest stages of the factor schedule solved wave-locally (each
                   wave solves whole independent subtrees, with no barrier between
                   those stages; for a split column, the stages above them are
                   solved part-locally, each workgroup with its own subtrees and
                   workgroup barriers): "auto" (default) chooses by a cost
                   estimate for columns split over workgroups and 0 otherwise; 0
                   solves every stage with all waves in step. Results do not
                   depend on it.
  column_pairs     True: split columns are worked two at a time by one set of
                   workgroups (half the workgroup size, twice the parts), which
                   stream the solve program once for both. Needs a split column,
                   workgroups of 512 or 1024 threads and the factor (with
                   solver="auto" a lattice circuit then uses the factor). Default False: on
                   SquareArray(512,512) it measured the same speed at W=16 and
                   1% faster at W=32. Results do not depend on it.
  solver           How the face system K is solved each step: "auto" (default)
                   takes the lattice solve for a rectangular lattice with one
                   junction weight per direction (no mutual inductance) -- a
                   sine transform of every lattice row (on the matrix cores for
                   a global face store), one tridiagonal solve per sine mode
                   (part by part in LDS for a split column), the inverse
                   transform -- and the staged Cholesky factor otherwise;
                   "lattice" requires the former, "cholesky" forces the latter.
                   Both are exact solves; results agree to rounding.
  max_resident_output_bytes
                   Device memory for stored steps (default 16 GiB). When the
                   stored theta and current exceed it, they stream to host
                   memory during the run through rings of output slots (each
                   launch's slots, with their vortex counts, are copied out
                   after it). Results do not depend on it.

There is no CPU fallback: without a HIP device ``compute()`` raises.
"""
from __future__ import annotations

import numpy as np

from .circuit import Circuit, DefaultCPR
from .solver_plan import STENCILS

__all__ = ["TimeEvolutionProblem", "TimeEvolutionResult", "SeparableSource", "ThetaNotStored", "CurrentNotStored",
           "VoltageNotStored", "DataAtTimepointNotStored"]


class ThetaNotStored(Exception):
    pass


class CurrentNotStored(Exception):
    pass


class VoltageNotStored(Exception):
    pass


class DataAtTimepointNotStored(Exception):
    pass


class SeparableSource:
    """An input whose value at element i, column w and step t is
    base[i, w] * amplitude[w, t] (base broadcastable to (N, W), amplitude to
    (W, Nt)), e.g. a bias pattern times a drive waveform:

        SeparableSource(circuit.current_base(0), Idc + 0.5 * np.sin(0.3 * dt * np.arange(Nt)))

    It stands for the array base[:, :, None] * amplitude[None, :, :] with the
    reference's broadcasting (time_evolution.py:119-130) without forming it: the
    device evaluates the product, one multiplication, as numpy does."""

    def __init__(self, base, amplitude):
        b = np.asarray(base, dtype=np.float64)
        a = np.asarray(amplitude, dtype=np.float64)
        self.base = b[:, None] if b.ndim == 1 else b
        self.amplitude = a[None, :] if a.ndim == 1 else a
        if self.base.ndim != 2 or self.amplitude.ndim != 2:
            raise ValueError("SeparableSource: base (N,) or (N, W), amplitude (Nt,) or (W, Nt)")
        if self.base.shape[1] not in (1, self.amplitude.shape[0]) and self.amplitude.shape[0] != 1:
            raise ValueError("SeparableSource: base and amplitude disagree on W")
        self.shape = (self.base.shape[0], max(self.base.shape[1], self.amplitude.shape[0]),
                      self.amplitude.shape[1])

    def __getitem__(self, key):
        i, w, t = key
        return (self.base[:, :, None] * self.amplitude[None, :, :])[i, w, t] \
            if not isinstance(t, (int, np.integer)) else \
            (self.base * self.amplitude[:, t][None, :])[i, w]

    def __array__(self, dtype=None, copy=None):
        out = self.base[:, :, None] * self.amplitude[None, :, :]
        return out if dtype is None else out.astype(dtype)


def _prob_cnt(x):
    if isinstance(x, SeparableSource):
        return x.shape[1]
    s = np.array(x(0) if callable(x) else x).shape
    return s[1] if len(s) > 1 else 1


class TimeEvolutionProblem:
    """Time evolution of W problems on one circuit (time_evolution.py:20-149).

    Parameters are those of the reference: time_step (dt), time_step_count (Nt),
    current_phase_relation, external_flux (f, (Nf,W,Nt)-broadcastable or
    callable i -> (Nf,W)), current_sources (Is), voltage_sources (Vs), temperature
    (T) (each (Nj,W,Nt)-broadcastable or callable), store_time_steps (None, index
    array or (Nt,) mask), store_theta/voltage/current, config_at_minus_1..4,
    stencil_width (3, 4 or 5).
    """

    def __init__(self, circuit: Circuit, time_step=0.05, time_step_count=1000,
                 current_phase_relation=None, external_flux=0.0, current_sources=0.0,
                 voltage_sources=0.0, temperature=0.0, store_time_steps=None, store_theta=True,
                 store_voltage=True, store_current=True, config_at_minus_1=None,
                 config_at_minus_2=None, config_at_minus_3=None, config_at_minus_4=None,
                 stencil_width=3, *, noise="philox", seed=None, steps_per_launch=None,
                 block_size=None, leaf_size=16, face_store="auto", parts=None, local_stages="auto",
                 column_pairs=False, max_resident_output_bytes=None, solver="auto"):
        self.circuit = circuit
        self.time_step = time_step
        self.time_step_count = time_step_count
        self.current_phase_relation = DefaultCPR() if current_phase_relation is None \
            else current_phase_relation
        self.problem_count = max(_prob_cnt(external_flux), _prob_cnt(current_sources),
                                 _prob_cnt(voltage_sources), _prob_cnt(temperature))
        Nj, Nf, W, Nt = circuit._Nj(), circuit._Nf(), self.problem_count, time_step_count
        self._f_is_timedep = self._is_timedep(external_flux)
        self._Is_is_timedep = self._is_timedep(current_sources)
        self._Vs_is_timedep = self._is_timedep(voltage_sources)
        self._T_is_timedep = self._is_timedep(temperature)
        def bc(x, n):
            if callable(x):
                return x
            if isinstance(x, SeparableSource):
                if x.shape[0] not in (1, n) or x.shape[2] not in (1, Nt):
                    raise ValueError("SeparableSource shape does not match the circuit or Nt")
                return x
            return np.broadcast_to(np.array(x), (n, W, Nt))
        # un-broadcast inputs: what is uploaded to the device (with zero strides)
        self._raw = {k: (v if callable(v) or isinstance(v, SeparableSource) else np.array(v)) for k, v in
                     (("f", external_flux), ("Is", current_sources), ("Vs", voltage_sources),
                      ("T", temperature))}
        self.external_flux = bc(external_flux, Nf)
        self.current_sources = bc(current_sources, Nj)
        self.voltage_sources = bc(voltage_sources, Nj)
        self.temperature = bc(temperature, Nj)
        self.store_time_steps = np.ones(Nt, dtype=bool)
        self.store_time_steps = self._to_time_point_mask(store_time_steps)
        self.store_theta = store_theta
        self.store_voltage = store_voltage
        self.store_current = store_current
        if not (store_theta or store_voltage or store_current):
            raise ValueError("No output is stored")
        if np.sum(self.store_time_steps) == 0:
            raise ValueError("No output is stored")
        self.stencil_width = stencil_width
        self.stencil = self._get_stencil(stencil_width)
        self.config_at_minus_1 = self._get_config(config_at_minus_1, np.zeros((Nj, W)), (Nj, W))
        self.config_at_minus_2 = self._get_config(config_at_minus_2, self.config_at_minus_1, (Nj, W))
        if stencil_width >= 4:
            self.config_at_minus_3 = self._get_config(config_at_minus_3, self.config_at_minus_2, (Nj, W))
        if stencil_width >= 5:
            self.config_at_minus_4 = self._get_config(config_at_minus_4, self.config_at_minus_3, (Nj, W))
        if noise not in ("philox", "reference"):
            raise ValueError("noise must be 'philox' or 'reference'")
        self.noise = noise
        self.seed = seed
        self.steps_per_launch = steps_per_launch
        self.block_size = block_size
        self.leaf_size = leaf_size
        self.face_store = face_store
        self.parts = parts
        self.local_stages = local_stages
        self.column_pairs = bool(column_pairs)
        self.max_resident_output_bytes = max_resident_output_bytes
        if solver not in ("auto", "lattice", "cholesky"):
            raise ValueError("solver must be 'auto', 'lattice' or 'cholesky'")
        self.solver = solver

    # ------------------------------------------------------------- getters
    def get_problem_count(self):
        return self.problem_count

    def get_static_problem(self, vortex_configuration, problem_nr=0, time_step=0):
        """StaticProblem with this problem's sources at one column and step
        (time_evolution.py:151-171)."""
        from .static_problem import StaticProblem
        return StaticProblem(self.circuit, current_sources=self._Is(time_step)[:, problem_nr].copy(),
                             external_flux=self._f(time_step)[:, problem_nr].copy(),
                             vortex_configuration=vortex_configuration,
                             current_phase_relation=self.current_phase_relation)

    def get_circuit(self) -> Circuit:
        return self.circuit

    def get_time_step(self):
        return self.time_step

    def get_time_step_count(self):
        return self.time_step_count

    def get_current_phase_relation(self):
        return self.current_phase_relation

    def get_phase_zone(self):
        return 0

    def get_external_flux(self):
        return self.external_flux

    def get_current_sources(self):
        return self.current_sources

    def get_voltage_sources(self):
        return self.voltage_sources

    def get_temperature(self):
        return self.temperature

    def get_store_time_steps(self):
        return self.store_time_steps

    def get_store_theta(self):
        return self.store_theta

    def get_store_voltage(self):
        return self.store_voltage

    def get_store_current(self):
        return self.store_current

    def get_time(self):
        return np.arange(self._Nt(), dtype=np.double) * self._dt()

    def get_net_sourced_current(self, time_step):
        M = self.get_circuit().get_cut_matrix()
        return 0.5 * np.sum(np.abs(M @ self._Is(time_step)), axis=0)

    def get_node_current_sources(self, time_step):
        return self.get_circuit().get_cut_matrix() @ self._Is(time_step)

    def compute(self) -> "TimeEvolutionResult":
        """Run the time evolution on the GPU (time_evolution.py:303-307)."""
        return time_evolution(self)

    def __str__(self):
        return (f"time evolution problem: {self.time_step_count} steps of {self.time_step}, "
                f"W={self.problem_count}, circuit {self.circuit}")

    # ------------------------------------------------------------- helpers
    def _Nt(self):
        return self.time_step_count

    def _Nt_s(self):
        return int(np.sum(self.store_time_steps))

    def _dt(self):
        return self.time_step

    @staticmethod
    def _get_config(cur, prev, shape):
        cur = prev.copy() if cur is None else cur
        if hasattr(cur, "get_theta"):
            cur = cur.get_theta()
        return np.asarray(cur, dtype=np.double).reshape(shape)

    @staticmethod
    def _is_timedep(x):
        if callable(x):
            return True
        if isinstance(x, SeparableSource):
            return x.shape[2] > 1
        s = np.array(x).shape
        return len(s) > 0 and s[-1] > 1

    def _slice(self, x, n, t):
        if callable(x):
            return np.broadcast_to(x(t), (n, self.problem_count))
        return x[:, :, t]

    def _f(self, t):
        return self._slice(self.external_flux, self.circuit._Nf(), t)

    def _Is(self, t):
        return self._slice(self.current_sources, self.circuit._Nj(), t)

    def _Vs(self, t):
        return self._slice(self.voltage_sources, self.circuit._Nj(), t)

    def _T(self, t):
        return self._slice(self.temperature, self.circuit._Nj(), t)

    def _cp(self, theta):
        return self.current_phase_relation.eval(self.circuit._Ic()[:, None], theta)

    def _dcp(self, theta):
        return self.current_phase_relation.d_eval(self.circuit._Ic()[:, None], theta)

    def _icp(self, theta):
        return self.current_phase_relation.i_eval(self.circuit._Ic()[:, None], theta)

    def _to_time_point_mask(self, time_points):
        if time_points is None:
            time_points = self.store_time_steps
        time_points = np.array(time_points)
        if time_points.dtype != bool:
            try:
                x = np.zeros(self._Nt(), dtype=bool)
                x[time_points] = True
                time_points = x
            except (IndexError, TypeError, ValueError):
                raise ValueError("Invalid store_time_steps; must be None, mask, slice or index array")
        return time_points

    def _get_stencil(self, width):
        if width not in STENCILS:
            raise ValueError(f"stencil width must be 3, 4 or 5 (equals {width})")
        return STENCILS[width]


def _apply_derivative(x, index, stencil, dt):
    acc = stencil[0] * x[:, :, index] + stencil[1] * x[:, :, index - 1]
    if len(stencil) >= 4:
        acc = acc + stencil[2] * x[:, :, index - 2]
    if len(stencil) >= 5:
        acc = acc + stencil[3] * x[:, :, index - 3]
    return acc / dt


def _masks(problem: TimeEvolutionProblem):
    """Steps whose theta and current the core must keep: the stored steps plus,
    for voltages, the stencil's earlier steps (time_evolution.py:422-440)."""
    Nt = problem._Nt()
    th_mask = problem.store_time_steps.copy() if (problem.store_theta or problem.store_voltage) \
        else np.zeros(Nt, dtype=bool)
    I_mask = problem.store_time_steps.copy() if (problem.store_current or problem.store_voltage) \
        else np.zeros(Nt, dtype=bool)
    Vth, VI = th_mask.copy(), I_mask.copy()
    t_ids = np.flatnonzero(problem.store_time_steps)
    Nj = problem.circuit._Nj()
    off = problem.stencil_width - 1
    if problem.store_voltage and len(t_ids):
        ids = (t_ids[:, None] - np.arange(off)).ravel()
        ids = ids[(ids >= 0) & (ids < Nt)]
        Vth[ids] = True
        if problem.circuit._has_inductance():
            VI[ids] = True
    return th_mask, I_mask, Vth, VI


def _finish(problem: TimeEvolutionProblem, masks, th_out, I_out, n_out):
    """Voltages by the stencil derivative (plus L dI/dt), helper steps and the
    prefix removed (time_evolution.py:441-458)."""
    th_mask, I_mask, Vth, VI = masks
    Nj = problem.circuit._Nj()
    off = problem.stencil_width - 1
    V_out = None
    if problem.store_voltage:
        ts = np.flatnonzero(problem.store_time_steps[Vth])
        V_out = _apply_derivative(th_out, ts + off, problem.stencil[0], problem._dt())
        if problem.circuit._has_inductance():
            V_ind = _apply_derivative(I_out, ts + off, problem.stencil[0], problem._dt())
            V_out += (problem.circuit.get_inductance() @ V_ind.reshape((Nj, -1))).reshape(V_out.shape)
        drop_th = np.flatnonzero((Vth & ~th_mask)[Vth])
        th_out = np.delete(th_out, drop_th + off, axis=2)
        I_out = np.delete(I_out, np.flatnonzero((VI & ~I_mask)[VI]) + off, axis=2)
        if n_out is not None:
            n_out = np.delete(n_out, drop_th, axis=2)
    th_out = th_out[:, :, off:]
    I_out = I_out[:, :, off:]
    return TimeEvolutionResult(problem, th_out if problem.store_theta else None,
                               I_out if problem.store_current else None,
                               V_out if problem.store_voltage else None,
                               vortex=n_out if problem.store_theta else None)


def time_evolution(problem: TimeEvolutionProblem):
    """Storage masks, GPU core, voltage post-processing (time_evolution.py:422-458)."""
    masks = _masks(problem)
    th_out, I_out, n_out = _gpu_core(problem, masks[2], masks[3])
    return _finish(problem, masks, th_out, I_out, n_out)


def _gpu_core(problem: TimeEvolutionProblem, th_mask, I_mask):
    """Replacement of time_evolution_core (time_evolution.py:461-582): returns
    (th_out, I_out, n_out) in the reference layout (Nj, W, prefix + stored) plus the
    vortex counts of the stored theta."""
    from .engine import Engine
    eng = Engine(problem, th_mask, I_mask)
    try:
        eng.advance(problem._Nt())
        return eng.results()
    finally:
        eng.close()


class TimeEvolutionResult:
    """Stored theta, current and voltage of a time evolution (time_evolution.py:597-1067)."""

    def __init__(self, problem: TimeEvolutionProblem, theta, current, voltage, vortex=None):
        self.problem = problem
        Nj, W, Nt_s = problem.circuit._Nj(), problem.get_problem_count(), problem._Nt_s()
        for name, val, flag in (("theta", theta, problem.store_theta),
                                ("current", current, problem.store_current),
                                ("voltage", voltage, problem.store_voltage)):
            if flag and val.shape != (Nj, W, Nt_s):
                raise ValueError(f"{name} must have shape {(Nj, W, Nt_s)}; has shape {val.shape}")
        self.theta = theta if problem.store_theta else None
        self.current = current if problem.store_current else None
        self.voltage = voltage if problem.store_voltage else None
        self._vortex = vortex
        s = problem.store_time_steps.astype(int)
        self.time_point_indices = np.cumsum(s) - s
        self.animation = None

    def _time_point_index(self, time_points):
        if time_points is None:
            time_points = self.problem.store_time_steps
        if not np.all(self.problem.store_time_steps[time_points]):
            raise DataAtTimepointNotStored("Queried a timepoint that is not stored during time "
                                           "evolution simulation.")
        return self.time_point_indices[time_points]

    def _th(self, tp):
        if self.theta is None:
            raise ThetaNotStored("Cannot query theta; quantity is not stored during time "
                                 "evolution simulation.")
        return self.theta[:, :, self._time_point_index(tp)]

    def _V(self, tp):
        if self.voltage is None:
            raise VoltageNotStored("Cannot query voltage; quantity is not stored during time "
                                   "evolution simulation.")
        return self.voltage[:, :, self._time_point_index(tp)]

    def _I(self, tp):
        if self.current is None:
            raise CurrentNotStored("Cannot query current; quantity is not stored during time "
                                   "evolution simulation.")
        return self.current[:, :, self._time_point_index(tp)]

    def _select(self, select_time_points, N, func, is_zero=False):
        tps = np.flatnonzero(self.problem._to_time_point_mask(select_time_points))
        out = np.zeros((N, self.get_problem_count(), len(tps)), dtype=np.double)
        if is_zero:
            return out
        for i, tp in enumerate(tps):
            out[:, :, i] = func(tp)
        return out

    def get_problem_count(self):
        return self.problem.get_problem_count()

    def get_circuit(self) -> Circuit:
        return self.problem.get_circuit()

    def get_theta(self, select_time_points=None):
        return self._select(select_time_points, self.get_circuit()._Nj(), self._th)

    def get_voltage(self, select_time_points=None):
        return self._select(select_time_points, self.get_circuit()._Nj(), self._V)

    def get_current(self, select_time_points=None):
        return self._select(select_time_points, self.get_circuit()._Nj(), self._I)

    def get_vortex_configuration(self, select_time_points=None):
        """n = -A round(theta / 2pi) (time_evolution.py:734-755), counted on the GPU for
        every stored step."""
        if self.theta is None:
            raise ThetaNotStored("Cannot compute n; requires theta to be stored in "
                                 "TimeEvolutionConfig")
        tps = np.flatnonzero(self.problem._to_time_point_mask(select_time_points))
        idx = self._time_point_index(tps)
        if self._vortex is not None and self._vortex.shape[0] == self.get_circuit()._Nf():
            return self._vortex[:, :, idx].astype(int)
        A = self.get_circuit().get_cycle_matrix()
        return self._select(select_time_points, self.get_circuit()._Nf(),
                            lambda tp: -A @ np.round(self._th(tp) / (2.0 * np.pi))).astype(int)

    # Derived quantities (time_evolution.py:692-983) are evaluated on the
    # device for all selected stored steps at once (result_queries.py); the
    # exceptions are the reference's.
    def _queries(self):
        if getattr(self, "_dq", None) is None:
            from .result_queries import DeviceQueries
            self._dq = DeviceQueries(self)
        return self._dq

    def _stored(self, name, exc, message):
        arr = getattr(self, name)
        if arr is None:
            raise exc(message)
        return arr

    def get_phase(self, select_time_points=None):
        th = self._stored("theta", ThetaNotStored, "Cannot compute phi; requires theta to be stored in "
                                                   "TimeEvolutionConfig")
        q = self._queries()
        return q.select(self.get_circuit()._Nn(), lambda tps, idx: q.phase(th, tps, idx), select_time_points)

    def get_potential(self, select_time_points=None):
        V = self._stored("voltage", VoltageNotStored, "Cannot compute electric potential U; requires voltage "
                                                      "to be stored in TimeEvolutionConfig")
        q = self._queries()
        return q.select(self.get_circuit()._Nn(), lambda tps, idx: q.potential(V, tps, idx), select_time_points)

    def get_josephson_energy(self, select_time_points=None):
        th = self._stored("theta", ThetaNotStored, "Cannot compute Josephson energy EJ; requires theta to be "
                                                   "stored in TimeEvolutionConfig")
        q = self._queries()
        return q.select(self.get_circuit()._Nj(), lambda tps, idx: q.josephson_energy(th, tps, idx),
                        select_time_points)

    def get_supercurrent(self, select_time_points=None):
        th = self._stored("theta", ThetaNotStored, "Cannot compute supercurrent Isup; requires theta to be "
                                                   "stored in TimeEvolutionConfig")
        q = self._queries()
        return q.select(self.get_circuit()._Nj(), lambda tps, idx: q.supercurrent(th, tps, idx),
                        select_time_points)

    def get_cycle_current(self, select_time_points=None):
        I = self._stored("current", CurrentNotStored, "Cannot compute cycle-current J; requires current to be "
                                                      "stored in TimeEvolutionConfig")
        q = self._queries()
        return q.select(self.get_circuit()._Nf(), lambda tps, idx: q.cycle_current(I, tps, idx),
                        select_time_points)

    def get_flux(self, select_time_points=None):
        I = self._stored("current", CurrentNotStored, "Cannot compute magnetic flux; requires current to be "
                                                      "stored in TimeEvolutionConfig")
        q = self._queries()
        return q.select(self.get_circuit()._Nf(), lambda tps, idx: q.flux(I, tps, idx), select_time_points)

    def get_magnetic_energy(self, select_time_points=None):
        c = self.get_circuit()
        if not c._has_inductance():          # zeros, stored or not (time_evolution.py:885-888)
            return self._select(select_time_points, c._Nj(), None, is_zero=True)
        I = self._stored("current", CurrentNotStored, "Cannot compute magnetic energy EM; requires current to "
                                                      "be stored in TimeEvolutionConfig")
        q = self._queries()
        return q.select(c._Nj(), lambda tps, idx: q.magnetic_energy(I, tps, idx), select_time_points)

    def get_capacitive_energy(self, select_time_points=None):
        c = self.get_circuit()
        if not c._has_capacitance():         # zeros, stored or not (time_evolution.py:957-960)
            return self._select(select_time_points, c._Nj(), None, is_zero=True)
        V = self._stored("voltage", VoltageNotStored, "Cannot compute capacitive energy EC; requires voltage "
                                                      "to be stored in TimeEvolutionConfig")
        q = self._queries()
        return q.select(c._Nj(), lambda tps, idx: q.capacitive_energy(V, tps, idx), select_time_points)

    def get_energy(self, select_time_points=None):
        return (self.get_josephson_energy(select_time_points) +
                self.get_magnetic_energy(select_time_points) +
                self.get_capacitive_energy(select_time_points))

    def plot(self, *args, **kwargs):
        raise NotImplementedError("visualisation is not part of pyjjasim_amd")

    def animate(self, *args, **kwargs):
        raise NotImplementedError("visualisation is not part of pyjjasim_amd")

    def __str__(self):
        shapes = ", ".join(f"{n}{v.shape}" for n, v in (("th", self.theta), ("I", self.current),
                                                       ("V", self.voltage)) if v is not None)
        return f"time evolution configuration: ({shapes})"
