"""Static (stationary) configurations by Newton's method on the device.

Drop-in for the reference's StaticProblem / StaticConfiguration / NewtonIterInfo
(static_problem.py:88-1107) as far as the time evolution needs them: a static
state as initial condition (TimeEvolutionProblem.get_static_problem,
time_evolution.py:151-171), the London approximation as the Newton start
(static_problem.py:550-557, 1385-1399), and the final static solve of
AnnealingProblem (time_evolution.py:1183-1191). The parameter maximisation
algorithms (static_problem.py:609-781, 1198-1380) are not part of this build.

Device form of each reference step:
  * Newton step (static_problem.py:1400-1515): q = Ic cos(theta), S = L +
    diag(1/q), y = (I - Is)/q, j = (A S A^T)^-1 (A (theta - y - L Is) + 2 pi
    (f - z)), theta -= y + A^T j / q. The reference factorises A S A^T with
    SuperLU at every iteration (Asq_solve_sandwich, josephson_circuit.py:
    153-174). Here A S A^T is applied as incidence gathers and solved by
    preconditioned MINRES (Paige and Saunders, 1975; it takes the symmetric
    indefinite systems of unstable junctions, q < 0), preconditioned by the
    London matrix A (diag(1/Ic) + L) A^T, whose Cholesky factor is computed once
    and applied by libjja's device solve (jja_solve). When S is a multiple of
    the identity, the reference's shortcut A A^T x = b / S_0 is taken, with the
    device solve of A A^T. The iteration's control (residual norms, stopping
    rules, status) is the reference's, on scalars read back each iteration.
  * London approximation: theta = diag(1/Ic) (A^T (A (diag(1/Ic) + L) A^T)^-1
    (2 pi (n - f) - A (diag(1/Ic) + L) Is) + Is), with the reference's special
    cases (_AiIcpLA_solve, josephson_circuit.py:176-189).
  * Phase-zone change: the integral cycle-space solve of the graph
    (embedded_graph.py:1041-1090), host integer work on the dual graph.
  * Stability (static_problem.py:1519-1588): -J = m diag(q) m^T + A L A^T
    terms must be positive definite; checked by the native sparse Cholesky
    (positive definite: it succeeds).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import scipy.sparse
import scipy.sparse.linalg

from . import _native
from .circuit import Circuit, DefaultCPR
from .graph import EmbeddedGraph

__all__ = ["StaticProblem", "StaticConfiguration", "NewtonIterInfo", "DEF_TOL", "DEF_NEWTON_MAXITER",
           "london_approximation", "static_compute", "get_kirchhoff_error", "get_winding_error",
           "principle_value", "change_phase_zone", "compute_stability"]

DEF_TOL = 1E-10
DEF_NEWTON_MAXITER = 30


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("pyjjasim_amd: static solves run on a HIP (ROCm) device; none is visible")
    return torch


class NewtonIterInfo:
    """Information about one Newton iteration (static_problem.py:88-201)."""

    def __init__(self, tol, maxiter):
        self.start_time = time.perf_counter()
        self.tol = tol
        self.maxiter = maxiter
        self.iteration = 0
        self.error = np.zeros(self.maxiter + 1, dtype=np.double)
        self.has_converged = False
        self.is_target_n = np.zeros(self.maxiter + 1, dtype=int)
        self.runtime = 0.0

    def get_max_iter(self):
        return self.maxiter

    def get_tol(self):
        return self.tol

    def get_status(self):
        """0: converged; 1: diverged before maxiter; 2: reached maxiter."""
        return int(not self.found_target_solution()) + 2 * int(self.iteration >= self.maxiter)

    def get_is_target_vortex_configuration(self):
        return self.is_target_n[:(self.get_number_of_iterations() + 1)]

    def found_target_solution(self):
        return self.has_converged and self.is_target_n[self._get_iteration()]

    def get_number_of_iterations(self):
        return self._get_iteration()

    def get_residual(self):
        return self.error[:(self.get_number_of_iterations() + 1)]

    def get_runtime(self):
        return self.runtime

    def plot_residuals(self):
        raise NotImplementedError("visualisation is not part of pyjjasim_amd")

    def __str__(self):
        out = f"newton iteration info: (tol={self.get_tol()}, maxiter={self.get_max_iter()})\n\t"
        out += f"status: {self.get_status()}"
        out += {0: " (converged)\n\t", 1: " (diverged)\n\t",
                2: " (indeterminate; reached max_iter without converging or diverging)\n\t"}[self.get_status()]
        out += f"number of iterations: {self.get_number_of_iterations()}\n\t"
        out += f"residual: {self.get_residual()}\n\t"
        out += f"runtime (in sec): {self.get_runtime()}"
        return out

    def _set(self, error, is_target_n_v):
        self.has_converged = error < self.tol
        self.is_target_n[self.iteration] = is_target_n_v
        self.error[self.iteration] = error
        self.runtime += time.perf_counter() - self.start_time
        self.start_time = time.perf_counter()
        self.iteration += 1
        return self

    def _get_iteration(self):
        return max(0, self.iteration - 1)


# ----------------------------------------------------------------- device
class _StaticDevice:
    """Per-circuit device operators of the static solves: incidence gathers A,
    A^T, M, the inductance L, the solvers of A A^T and of the London matrix,
    and the norms of A and M (computed once, as the reference caches them)."""

    def __init__(self, circuit):
        from .result_queries import PaddedRows
        self.torch = torch = _torch()
        self.circuit = circuit
        A = circuit.get_cycle_matrix()
        self.A = PaddedRows(A)
        self.AT = PaddedRows(A.T)
        self.M = PaddedRows(circuit.get_cut_matrix())
        self.L = PaddedRows(circuit._L())
        self.Ic = torch.from_numpy(np.ascontiguousarray(circuit._Ic(), dtype=np.float64)).cuda()
        self._cycle = None
        self._london = None
        self._norms = None

    def dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()

    def cycle(self):
        if self._cycle is None:
            from .device_linalg import cycle_solver
            self._cycle = cycle_solver(self.circuit)
        return self._cycle

    def london(self):
        """Solver of A (diag(iIc) + L) A^T, iIc = 1/Ic where |Ic| > 1e-12."""
        if self._london is None:
            from .device_linalg import DeviceSolver
            c = self.circuit
            A = c.get_cycle_matrix()
            cx, cy = c.get_face_centroids()
            K = A @ (scipy.sparse.diags(_inv_ic(c._Ic())) + c._L()) @ A.T
            self._london = DeviceSolver(K.tocsc(), np.stack([cx, cy], axis=1))
        return self._london

    def norms(self):
        """||A||_2 and ||M||_2 (josephson_circuit.py:762-780): the square roots of
        the largest eigenvalues of A A^T and M M^T."""
        if self._norms is None:
            out = []
            for B in (self.circuit.get_cycle_matrix(), self.circuit.get_cut_matrix()):
                if B.shape[0] == 1:
                    out.append(float(np.sqrt((B @ B.T).todense()[0, 0])))
                else:
                    out.append(float(np.sqrt(scipy.sparse.linalg.eigsh((B @ B.T).astype(np.double), 1,
                                                                       maxiter=1000, which="LA")[0][0])))
            self._norms = tuple(out)
        return self._norms

    def close(self):
        for s in (self._cycle, self._london):
            if s is not None:
                s.close()


def _device(circuit):
    d = getattr(circuit, "_static_device", None)
    if d is None:
        d = _StaticDevice(circuit)
        circuit._static_device = d
    return d


def _inv_ic(Ic):
    Ic = np.array(Ic, dtype=np.double)
    big = np.abs(Ic) > 1E-12
    Ic[big] = Ic[big] ** -1
    return Ic


def _identical(x):
    return bool(np.allclose(x, x[0]))


def _minres(matvec, psolve, b, rtol=1e-14, maxiter=1000):
    """Preconditioned MINRES (Paige and Saunders, SIAM J. Numer. Anal. 12, 1975)
    for a symmetric, possibly indefinite operator with a symmetric positive
    definite preconditioner, on device tensors. Stops when the preconditioned
    residual estimate falls below rtol times its initial value."""
    torch = _torch()
    x = torch.zeros_like(b)
    r1 = b.clone()
    y = psolve(r1)
    beta1 = float(torch.dot(r1, y))
    if beta1 <= 0:
        return x
    beta1 = beta1 ** 0.5
    oldb, beta, dbar, epsln, phibar = 0.0, beta1, 0.0, 0.0, beta1
    cs, sn = -1.0, 0.0
    w = torch.zeros_like(b)
    w2 = torch.zeros_like(b)
    r2 = r1.clone()
    eps = np.finfo(float).eps
    for itn in range(maxiter):
        v = y / beta
        y = matvec(v)
        if itn >= 1:
            y = y - (beta / oldb) * r1
        alfa = float(torch.dot(v, y))
        y = y - (alfa / beta) * r2
        r1, r2 = r2, y
        y = psolve(r2)
        oldb = beta
        beta2 = float(torch.dot(r2, y))
        beta = max(beta2, 0.0) ** 0.5
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max((gbar * gbar + beta * beta) ** 0.5, eps)
        cs, sn = gbar / gamma, beta / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1, w2 = w2, w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        if phibar <= rtol * beta1 or beta == 0.0:
            break
    return x


def _asq_solve_sandwich(dev, b, q):
    """(A S A^T) x = b with S = L + diag(1/q) (josephson_circuit.py:153-174)."""
    torch = dev.torch
    c = dev.circuit
    L = c._L()
    if c._has_only_self_inductance() or L.nnz == 0:
        Sd = L.diagonal() + 1.0 / q.cpu().numpy()
        if np.count_nonzero(Sd) == len(Sd) and np.allclose(Sd[0], Sd, rtol=1E-14):
            return dev.cycle().solve(b) / Sd[0]
    iq = 1.0 / q

    def matvec(v):
        u = dev.AT(v[:, None])[:, 0]
        return dev.A((dev.L(u[:, None])[:, 0] + iq * u)[:, None])[:, 0]
    return _minres(matvec, dev.london().solve, b)


# ------------------------------------------------------------ algorithms
def get_kirchhoff_error(circuit: Circuit, I, Is, precomputed_Is_norm=None):
    """Normalised residual of M (I - Is) = 0 (static_problem.py:1108-1117)."""
    torch = _torch()
    dev = _device(circuit)
    I, Is = _as_dev(dev, I), _as_dev(dev, Is)
    if precomputed_Is_norm is None:
        precomputed_Is_norm = float(torch.linalg.norm(Is))
    b = dev.M((I - Is)[:, None])[:, 0]
    normalizer = dev.norms()[1] * (precomputed_Is_norm + float(torch.linalg.norm(I)))
    return np.finfo(float).eps if abs(normalizer) < 1E-20 else float(torch.linalg.norm(b)) / normalizer


def get_winding_error(circuit: Circuit, th, I, df):
    """Normalised residual of A (theta + L I) + df = 0 (static_problem.py:1119-1130)."""
    torch = _torch()
    dev = _device(circuit)
    th, I, df = _as_dev(dev, th), _as_dev(dev, I), _as_dev(dev, df)

    def norm(x):
        return float(torch.linalg.norm(x)) / np.sqrt(len(x))
    LI = dev.L(I[:, None])[:, 0]
    normalizer = dev.norms()[0] * (norm(th) + norm(LI)) + norm(df)
    r = df + dev.A((th + LI)[:, None])[:, 0]
    return np.finfo(float).eps if abs(normalizer) < 1E-20 else norm(r) / normalizer


def _as_dev(dev, x):
    torch = dev.torch
    if isinstance(x, torch.Tensor):
        return x
    return dev.dev(x)


def principle_value(theta):
    """theta in [-pi, pi) (static_problem.py:1132-1136)."""
    return theta - 2 * np.pi * np.round(theta / (2 * np.pi))


def change_phase_zone(circuit: Circuit, theta, z_old, z_new):
    """theta + 2 pi Z with A Z = z_new - z_old integral (static_problem.py:1145-1172)."""
    if np.all(z_new == z_old):
        return theta
    return theta + circuit._A_solve(np.broadcast_to(z_new - z_old, (circuit._Nf(),)).copy()) * 2.0 * np.pi


def london_approximation(circuit: Circuit, f, n, Is):
    """Linearised (London) state (static_problem.py:1385-1399), on the device."""
    Nj = circuit.junction_count()
    Icv = circuit._Ic()
    if _identical(Icv) and np.abs(Icv[0]) < 1E-12:
        return np.zeros(Nj, dtype=np.double)
    dev = _device(circuit)
    Nf = circuit._Nf()
    iIc = dev.dev(_inv_ic(Icv))
    df = dev.dev(2 * np.pi * np.broadcast_to(np.asarray(n, dtype=np.double) - f, (Nf,)))
    Isc = dev.dev(np.broadcast_to(Is, (Nj,)))
    rhs = df - dev.A((iIc * Isc + dev.L(Isc[:, None])[:, 0])[:, None])[:, 0]
    L = circuit._L()
    if _identical(Icv) and (L.nnz == 0 or (circuit._has_only_self_inductance() and _identical(L.diagonal()))):
        const = 1 / Icv[0] + (L.diagonal()[0] if L.nnz else 0.0)
        sol = dev.cycle().solve(rhs) / const
    else:
        sol = dev.london().solve(rhs)
    theta = iIc * (dev.AT(sol[:, None])[:, 0] + Isc)
    return theta.cpu().numpy()


def static_compute(circuit: Circuit, theta0, Is, f, n, z=0, cp=None, tol=DEF_TOL, maxiter=DEF_NEWTON_MAXITER,
                   stop_as_residual_increases=True, stop_if_not_target_n=False):
    """Newton iteration for a stationary state (static_problem.py:1400-1515), on
    the device. Returns (theta (Nj,), status, NewtonIterInfo)."""
    if cp is not None and not isinstance(cp, DefaultCPR):
        raise ValueError("the device Newton solve evaluates the default current-phase relation "
                         "Ic*sin(theta) only; got a custom CurrentPhaseRelation")
    torch = _torch()
    info = NewtonIterInfo(tol, maxiter)
    Nj, Nf = circuit._Nj(), circuit._Nf()
    A = circuit.get_cycle_matrix()
    dev = _device(circuit)
    Ic = dev.Ic
    Is = np.ones((Nj,), dtype=np.double) * Is if np.array(Is).size == 1 else np.asarray(Is, dtype=np.double)
    f = np.ones((Nf,), dtype=np.double) * f if np.array(f).size == 1 else np.asarray(f, dtype=np.double)
    n = np.ones((Nf,), dtype=int) * n if np.array(n).size == 1 else np.asarray(n)
    target = np.broadcast_to(z - n, (Nf,))
    df = dev.dev(2 * np.pi * (f - z))
    Is_d = dev.dev(Is)
    theta = dev.dev(np.array(theta0, dtype=np.double))
    I = Ic * torch.sin(theta)
    LIs = dev.L(Is_d[:, None])[:, 0]
    Is_norm = float(torch.linalg.norm(Is_d))

    def errors(theta, I):
        e1 = get_kirchhoff_error(circuit, I, Is_d, precomputed_Is_norm=Is_norm)
        e2 = get_winding_error(circuit, theta, I, df)
        th = theta.cpu().numpy()
        return max(e1, e2), bool(np.all(A @ np.round(th / (2 * np.pi)).astype(int) == target))

    error, is_target_n = errors(theta, I)
    info._set(error, is_target_n)
    prev_error = np.inf
    iteration = 0
    while not (error < tol or (error > 0.5 and iteration > 5) or (stop_if_not_target_n and is_target_n) or
               (stop_as_residual_increases and error > prev_error) or iteration >= maxiter):
        q = Ic * torch.cos(theta)
        q = torch.where(q.abs() < 0.1 * tol, torch.full_like(q, 0.1 * tol), q)
        y = (I - Is_d) / q
        rhs = dev.A((theta - y - LIs)[:, None])[:, 0] + df
        j = _asq_solve_sandwich(dev, rhs, q)
        if not bool(torch.isfinite(j).all()):
            theta = theta + 10 ** 10
        else:
            theta = theta - (y + dev.AT(j[:, None])[:, 0] / q)
        I = Ic * torch.sin(theta)
        error, is_target_n = errors(theta, I)
        info._set(error, is_target_n)
        if iteration >= 3:
            prev_error = info.error[iteration - 3]
        iteration += 1
    return theta.cpu().numpy(), info.get_status(), info


def stability_scheme_0(circuit: Circuit, theta, cp=None):
    """J = -(m diag(q) m^T + A L A^T at the cycle block), m = [Mr; A L], with
    empty rows and columns removed (static_problem.py:1564-1588)."""
    Nj, Nnr = circuit._Nj(), circuit._Nnr()
    A, M, L = circuit.get_cycle_matrix(), circuit._Mr(), circuit._L()
    q = circuit._Ic() * np.cos(theta)
    AL = (A @ L @ A.T).tocoo()
    ALL = scipy.sparse.coo_matrix((AL.data, (AL.row + Nnr, AL.col + Nnr)), shape=(Nj, Nj)).tocsc()
    m = scipy.sparse.vstack([M, A @ L]).tocsc()
    J = -(m @ scipy.sparse.diags(q, 0) @ m.T + ALL)
    select = np.diff(J.indptr) != 0
    return J[select, :][:, select]


def compute_stability(circuit: Circuit, theta, cp=None):
    """0 if -J is positive definite (stable), 1 otherwise
    (static_problem.py:1519-1536): the native sparse Cholesky of -J succeeds
    exactly when it is."""
    from .solver_plan import cholesky
    X = (-stability_scheme_0(circuit, theta, cp)).tocsc()
    X.sum_duplicates()
    try:
        cholesky(X)
    except ValueError:
        return 1
    return 0


# ---------------------------------------------------------------- classes
class StaticProblem:
    """A static problem (static_problem.py:389-813): circuit, source currents Is
    (Nj,), external flux f (Nf,), target vortex configuration n (Nf,)."""

    def __init__(self, circuit: Circuit, current_sources=0.0, external_flux=0.0,
                 vortex_configuration=0, current_phase_relation=None):
        self.circuit = circuit
        self.current_sources = np.atleast_1d(current_sources)
        self.external_flux = np.atleast_1d(external_flux)
        self.vortex_configuration = np.atleast_1d(vortex_configuration)
        self.current_phase_relation = DefaultCPR() if current_phase_relation is None else current_phase_relation
        self.current_sources_norm = None

    def save(self, filename):
        """Sequential np.save records (static_problem.py:433-457)."""
        with open(filename, "wb") as fh:
            self.circuit._save_records(fh)
            np.save(fh, self.current_sources)
            np.save(fh, self.external_flux)
            np.save(fh, self.vortex_configuration)

    @staticmethod
    def load(filename) -> "StaticProblem":
        with open(filename, "rb") as fh:
            circuit = Circuit._load_records(fh)
            Is, f, n = np.load(fh), np.load(fh), np.load(fh)
        return StaticProblem(circuit, current_sources=Is, external_flux=f, vortex_configuration=n)

    def get_circuit(self) -> Circuit:
        return self.circuit

    def get_current_sources(self):
        return self.current_sources

    def get_external_flux(self):
        return self.external_flux

    def get_vortex_configuration(self):
        return self.vortex_configuration

    def get_current_phase_relation(self):
        return self.current_phase_relation

    def new_problem(self, current_sources=None, external_flux=None, vortex_configuration=None,
                    current_phase_relation=None) -> "StaticProblem":
        return StaticProblem(
            self.circuit, current_sources=self.current_sources if current_sources is None else current_sources,
            external_flux=self.external_flux if external_flux is None else external_flux,
            vortex_configuration=self.vortex_configuration if vortex_configuration is None else vortex_configuration,
            current_phase_relation=self.current_phase_relation if current_phase_relation is None
            else current_phase_relation)

    def get_phase_zone(self):
        return 0

    def get_net_sourced_current(self):
        M = self.get_circuit().get_cut_matrix()
        return 0.5 * np.sum(np.abs((M @ self._Is())), axis=0)

    def get_node_current_sources(self):
        M = self.get_circuit().get_cut_matrix()
        return M @ self.current_sources

    def approximate(self) -> "StaticConfiguration":
        """London approximation moved to phase zone 0 (static_problem.py:550-557)."""
        theta = london_approximation(self.circuit, self._f(), self._nt(), self._Is())
        theta = change_phase_zone(self.get_circuit(), theta, self._nt(), 0)
        return StaticConfiguration(self, theta)

    def compute(self, initial_guess=None, tol=DEF_TOL, maxiter=DEF_NEWTON_MAXITER,
                stop_as_residual_increases=True, stop_if_not_target_n=False):
        """Newton iteration from initial_guess (theta or a StaticConfiguration;
        None: the approximation); returns (config, status, NewtonIterInfo)
        (static_problem.py:559-607)."""
        if initial_guess is None:
            initial_guess = self.approximate()
        if isinstance(initial_guess, StaticConfiguration):
            initial_guess = initial_guess._th()
        initial_guess = np.array(initial_guess, dtype=np.double)
        theta, status, info = static_compute(self.get_circuit(), initial_guess, Is=self._Is(), f=self._f(),
                                             n=self._nt(), z=0, cp=self.current_phase_relation, tol=tol,
                                             maxiter=maxiter,
                                             stop_as_residual_increases=stop_as_residual_increases,
                                             stop_if_not_target_n=stop_if_not_target_n)
        return StaticConfiguration(self, theta), status, info

    def _not_here(self, *args, **kwargs):
        raise NotImplementedError("parameter maximisation (static_problem.py:609-781) is not part of "
                                  "pyjjasim_amd")

    compute_external_flux_bounds = _not_here
    compute_maximal_current = _not_here
    compute_stable_region = _not_here

    def __str__(self):
        return f"static problem: circuit {self.circuit}"

    def _Is(self):
        return self.current_sources

    def _f(self):
        return self.external_flux

    def _nt(self):
        return self.vortex_configuration

    def _cp(self, Ic, theta):
        return self.current_phase_relation.eval(Ic, theta)

    def _dcp(self, Ic, theta):
        return self.current_phase_relation.d_eval(Ic, theta)

    def _icp(self, Ic, theta):
        return self.current_phase_relation.i_eval(Ic, theta)

    def _Is_norm(self):
        if self.current_sources_norm is None:
            self.current_sources_norm = float(np.linalg.norm(np.broadcast_to(self.current_sources,
                                                                            (self.circuit._Nj(),))))
        return self.current_sources_norm


class StaticConfiguration:
    """A state theta (Nj,) of a static problem and its derived quantities
    (static_problem.py:815-1105); the quantities are evaluated on the device."""

    def __init__(self, problem: StaticProblem, theta):
        self.problem = problem
        self.theta = np.array(theta)
        if not self.theta.shape == (self.problem.get_circuit()._Nj(),):
            raise ValueError("theta must be of shape (Nj,)")

    def _dev(self):
        return _device(self.get_circuit())

    def get_circuit(self) -> Circuit:
        return self.problem.get_circuit()

    def get_problem(self) -> StaticProblem:
        return self.problem

    def get_phase(self):
        """Node phases, the last node grounded (static_problem.py:856-862)."""
        from .device_linalg import cut_solver
        dev = self._dev()
        torch = dev.torch
        c = self.get_circuit()
        if getattr(dev, "_cut", None) is None:
            dev._cut = cut_solver(c)
        b = dev.M(dev.dev(self.theta)[:, None])[:, 0]
        x = torch.cat((dev._cut.solve(b[:-1].contiguous()), torch.zeros_like(b[-1:])))
        return x.cpu().numpy()

    def get_theta(self):
        return self.theta

    def get_vortex_configuration(self):
        A = self.get_circuit().get_cycle_matrix()
        return - (A @ np.round(self._th() / (2.0 * np.pi))).astype(int)

    def _current_dev(self):
        dev = self._dev()
        return dev.Ic * dev.torch.sin(dev.dev(self.theta))

    def get_current(self):
        return self._current_dev().cpu().numpy()

    def get_cycle_current(self):
        dev = self._dev()
        d = self._current_dev() - dev.dev(np.broadcast_to(self.problem.current_sources, (self.theta.shape[0],)))
        return dev.cycle().solve(dev.A(d[:, None])[:, 0]).cpu().numpy()

    def get_flux(self):
        dev = self._dev()
        LI = dev.L(self._current_dev()[:, None])[:, 0]
        return self.problem.external_flux + (dev.A(LI[:, None])[:, 0] / (2 * np.pi)).cpu().numpy()

    def get_magnetic_energy(self):
        from .result_queries import PaddedRows
        dev = self._dev()
        if getattr(dev, "_Lhalf", None) is None:
            dev._Lhalf = PaddedRows(self.get_circuit()._L(), 0.5)
        return dev._Lhalf((self._current_dev() ** 2)[:, None])[:, 0].cpu().numpy()

    def get_josephson_energy(self):
        dev = self._dev()
        return (dev.Ic * (1.0 - dev.torch.cos(dev.dev(self.theta)))).cpu().numpy()

    def get_energy(self):
        return self.get_josephson_energy() + self.get_magnetic_energy()

    def satisfies_kirchhoff_rules(self, tol=DEF_TOL):
        return self.get_error_kirchhoff_rules() < tol

    def satisfies_winding_rules(self, tol=DEF_TOL):
        return self.get_error_winding_rules() < tol

    def satisfies_target_vortices(self):
        return np.all(self.get_vortex_configuration() == self.problem.get_vortex_configuration())

    def is_stable(self) -> int:
        """0: stable, 1: unstable (static_problem.py:938-952)."""
        return compute_stability(self.get_circuit(), self._th(), self.problem.get_current_phase_relation())

    def is_solution(self, tol=DEF_TOL):
        return self.satisfies_kirchhoff_rules(tol) & self.satisfies_winding_rules(tol)

    def is_target_solution(self, tol=DEF_TOL):
        return self.is_solution(tol=tol) & self.satisfies_target_vortices()

    def is_stable_target_solution(self, tol=DEF_TOL):
        return self.is_target_solution(tol=tol) & (self.is_stable() == 0)

    def get_error_kirchhoff_rules(self):
        return get_kirchhoff_error(self.get_circuit(), self._current_dev(),
                                   np.broadcast_to(self.problem._Is(), self.theta.shape),
                                   precomputed_Is_norm=self.problem._Is_norm())

    def get_error_winding_rules(self):
        f = np.broadcast_to(self.problem._f(), (self.get_circuit()._Nf(),))
        return get_winding_error(self.get_circuit(), self.theta, self._current_dev(), 2 * np.pi * f)

    def get_error(self):
        return self.get_error_kirchhoff_rules(), self.get_error_winding_rules()

    def plot(self, *args, **kwargs):
        raise NotImplementedError("visualisation is not part of pyjjasim_amd")

    def report(self):
        print("Kirchhoff rules error:    ", self.get_error_kirchhoff_rules())
        print("Path rules error:         ", self.get_error_winding_rules())
        print("is stable:                ", self.is_stable() == 0)
        print("is target vortex solution:", self.satisfies_target_vortices())

    def save(self, filename):
        """The problem's records followed by theta (the layout of
        static_problem.py:1036-1062)."""
        with open(filename, "wb") as fh:
            self.problem.circuit._save_records(fh)
            np.save(fh, self.problem.current_sources)
            np.save(fh, self.problem.external_flux)
            np.save(fh, self.problem.vortex_configuration)
            np.save(fh, self.theta)

    @staticmethod
    def load(filename) -> "StaticConfiguration":
        with open(filename, "rb") as fh:
            circuit = Circuit._load_records(fh)
            Is, f, n = np.load(fh), np.load(fh), np.load(fh)
            th = np.load(fh)
        prob = StaticProblem(circuit, current_sources=Is, external_flux=f, vortex_configuration=n)
        return StaticConfiguration(prob, th)

    def _th(self):
        return self.theta

    def __str__(self):
        return f"Static configuration with theta={self.theta}"
