"""Josephson circuits: per-junction RCSJ parameters on a planar embedded graph.

Host-side data provider for the time-evolution path. Mirrors the parts of the
reference's ``Circuit`` that the time evolution and its result queries use
(josephson_circuit.py:22-1020), plus the lattice constructors and the
``current_base`` bias helper.

All linear solves here (``Msq_solve``, ``Asq_solve``) run once to build inputs or
answer post-run queries; they use the same SuperLU calls as the reference so that
bias vectors built from them are bit-identical to the reference's.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse
import scipy.sparse.linalg
import scipy.spatial

from .graph import (EmbeddedGraph, EmbeddedSquareGraph, EmbeddedHoneycombGraph,
                    EmbeddedTriangularGraph)

__all__ = ["Circuit", "SquareArray", "HoneycombArray", "TriangularArray", "SQUID",
           "CurrentPhaseRelation", "DefaultCPR", "FourierCPR", "node_to_junction_current",
           "NoCurrentConservationError"]


class NoCurrentConservationError(Exception):
    pass


class CurrentPhaseRelation:
    """Current-phase relation Icp(Ic, theta) with derivative and integral
    (static_problem.py:34-75). ``func``, ``d_func`` and ``i_func`` are numpy ufunc-style
    callables of (Ic, theta)."""

    def __init__(self, func, d_func, i_func):
        self.func = func
        self.d_func = d_func
        self.i_func = i_func

    def eval(self, Ic, theta):
        return self.func(Ic, theta)

    def d_eval(self, Ic, theta):
        return self.d_func(Ic, theta)

    def i_eval(self, Ic, theta):
        return self.i_func(Ic, theta)


def _sin_cpr(Ic, th):
    return Ic * np.sin(th)


def _cos_cpr(Ic, th):
    return Ic * np.cos(th)


def _int_cpr(Ic, th):
    return Ic * (1.0 - np.cos(th))


class DefaultCPR(CurrentPhaseRelation):
    """Icp = Ic sin(theta) (static_problem.py:77-85). The only relation the GPU step
    kernel evaluates; it is recognised by type."""

    def __init__(self):
        super().__init__(_sin_cpr, _cos_cpr, _int_cpr)


class FourierCPR(CurrentPhaseRelation):
    """Icp = Ic * sum_k a_k sin(k theta), k = 1 .. n with n <= 8: a sine-series
    current-phase relation (higher harmonics, skewed relations), the family the
    GPU step kernel evaluates besides the default. Its derivative
    Ic sum_k k a_k cos(k theta) and integral Ic sum_k a_k (1 - cos(k theta)) / k
    complete it as a CurrentPhaseRelation (static_problem.py:34-75).
    FourierCPR([1.0]) is the default relation."""

    MAX_TERMS = 8

    def __init__(self, coefficients):
        a = np.atleast_1d(np.asarray(coefficients, dtype=np.float64))
        if a.ndim != 1 or not 1 <= len(a) <= self.MAX_TERMS:
            raise ValueError(f"FourierCPR takes 1 to {self.MAX_TERMS} sine coefficients")
        self.coefficients = a

        def func(Ic, th):
            acc = a[0] * np.sin(th)
            for k in range(2, len(a) + 1):
                acc = acc + a[k - 1] * np.sin(k * th)
            return Ic * acc

        def d_func(Ic, th):
            acc = a[0] * np.cos(th)
            for k in range(2, len(a) + 1):
                acc = acc + (k * a[k - 1]) * np.cos(k * th)
            return Ic * acc

        def i_func(Ic, th):
            acc = a[0] * (1.0 - np.cos(th))
            for k in range(2, len(a) + 1):
                acc = acc + (a[k - 1] / k) * (1.0 - np.cos(k * th))
            return Ic * acc
        super().__init__(func, d_func, i_func)


def device_cpr_terms(cpr):
    """(terms, coefficients) of a relation the step kernel evaluates: 0 terms for
    the default Ic sin(theta), the sine series of a FourierCPR; None otherwise."""
    if type(cpr) is DefaultCPR:
        return 0, np.zeros(FourierCPR.MAX_TERMS)
    if isinstance(cpr, FourierCPR):
        c = np.zeros(FourierCPR.MAX_TERMS)
        c[:len(cpr.coefficients)] = cpr.coefficients
        return len(cpr.coefficients), c
    return None


class Circuit:
    """Josephson circuit (josephson_circuit.py:22-91).

    Parameters
    ----------
    graph : EmbeddedGraph (planar embedding, single component)
    critical_current, resistance, capacitance : scalar or (Nj,) array
    inductance : scalar, (Nj,) array, or symmetric positive definite (Nj, Nj) matrix
    """

    def __init__(self, graph: EmbeddedGraph, critical_current=1.0, resistance=1.0,
                 capacitance=0.0, inductance=0.0, negative_Ic_allowed=False):
        self.graph = graph
        graph._assert_planar_embedding()
        graph._assert_single_component()
        self.negative_Ic_allowed = negative_Ic_allowed
        self.locator = None
        self.cut_matrix = graph.cut_space_matrix()
        self.cycle_matrix = graph.face_cycle_matrix()
        self.critical_current = None
        self.resistance = None
        self.capacitance = None
        self.inductance = None
        self._has_inductance_v = False
        self.set_resistance(resistance)
        self.set_critical_current(critical_current)
        self.set_capacitance(capacitance)
        self.set_inductance(inductance)
        self.cut_matrix_reduced = None
        self._Msq_factorized = None
        self._Asq_factorized = None

    def __getstate__(self):
        # cached factorisations are rebuilt on demand, not pickled
        state = dict(self.__dict__)
        state["_Msq_factorized"] = None
        state["_Asq_factorized"] = None
        return state

    # ------------------------------------------------------------- counts
    def _Nn(self):
        return self.graph.node_count()

    def _Nnr(self):
        return self._Nn() - 1

    def _Nj(self):
        return self.graph.edge_count()

    def _Nf(self):
        return self.graph.face_count()

    def node_count(self):
        return self._Nn()

    def junction_count(self):
        return self._Nj()

    def face_count(self):
        return self._Nf()

    # ---------------------------------------------------------- parameters
    def _Ic(self):
        return self.critical_current

    def _R(self):
        return self.resistance

    def _C(self):
        return self.capacitance

    def _L(self):
        return self.inductance

    @staticmethod
    def _prepare_junction_quantity(x, N, x_name="x"):
        try:
            return np.broadcast_to(x, (N,)).copy().astype(np.double)
        except ValueError:
            raise ValueError(x_name + " must be scalar or array of length equal to junction count")

    def get_critical_current(self):
        return self.critical_current

    def set_critical_current(self, Ic):
        self.critical_current = self._prepare_junction_quantity(Ic, self._Nj(), "Ic")
        if not self.negative_Ic_allowed and np.any(self.critical_current < 0):
            raise ValueError("Defined negative critical current. If intentional;"
                             "set negative_Ic_allowed field of circuit to True.")
        return self

    def get_resistance(self):
        return self.resistance

    def set_resistance(self, R):
        self.resistance = self._prepare_junction_quantity(R, self._Nj(), "R")
        if np.any(self.resistance <= 0.0):
            raise ValueError("All junctions must have a positive resistor")
        return self

    def get_capacitance(self):
        return self.capacitance

    def set_capacitance(self, C):
        self.capacitance = self._prepare_junction_quantity(C, self._Nj(), "C")
        if np.any(self.capacitance < 0.0):
            raise ValueError("Capacitance cannot be negative.")
        return self

    def get_inductance(self):
        return self.inductance

    def set_inductance(self, inductance):
        """Accepts scalar, (Nj,) (self inductance) or a symmetric positive definite
        (Nj, Nj) dense or sparse matrix (josephson_circuit.py:595-604, 718-756)."""
        self.inductance, pd, self._has_inductance_v = self._prepare_inductance(inductance,
                                                                               self._Nj())
        # a full matrix given as a dense array is written back dense by save(),
        # as the reference keeps and saves it (josephson_circuit.py:650-658)
        self._inductance_dense = (np.ndim(inductance) == 2 and not scipy.sparse.issparse(inductance))
        if not pd:
            raise ValueError("Inductance matrix not positive definite")
        return self

    @staticmethod
    def _prepare_inductance(L, N):
        eps = 10 * np.finfo(float).eps
        if not hasattr(L, "ndim"):
            L = np.array(L)
        if L.ndim <= 1:
            d = Circuit._prepare_junction_quantity(L, N, "L")
            return scipy.sparse.diags(d, 0).tocsc(), bool(np.all(d > -eps)), bool(np.any(d != 0.0))
        if L.shape != (N, N):
            raise ValueError("L must be scalar, (Nj,) array or (Nj, Nj) matrix")
        # For a full (Nj, Nj) matrix the reference's third return value (the
        # "has inductance" flag) is the matrix's is-zero test
        # (josephson_circuit.py:730-756): a non-zero mutual-inductance matrix
        # enters K but adds no L dI/dt term to the voltage. Kept as is so that
        # results match the reference on the same inputs.
        if scipy.sparse.issparse(L):
            if (L - L.T).nnz != 0:
                raise ValueError("inductance matrix must be symmetric")
            L = scipy.sparse.csc_matrix(L)
            if L.nnz == 0:
                return L, True, True
            return L, _is_positive_definite_sparse(L), False
        L = np.asarray(L, dtype=np.double)
        if not np.all(L == L.T):
            raise ValueError("inductance matrix must be symmetric")
        if np.all(L == 0):
            return scipy.sparse.csc_matrix(L), True, True
        try:
            np.linalg.cholesky(L)
            pd = True
        except np.linalg.LinAlgError:
            pd = False
        return scipy.sparse.csc_matrix(L), pd, False

    def _has_inductance(self):
        return self._has_inductance_v

    def _has_capacitance(self):
        return bool(np.any(self.capacitance > 0))

    def _has_only_self_inductance(self):
        L = self.inductance
        return L.nnz == np.count_nonzero(L.diagonal())

    # ---------------------------------------------------------- topology
    def get_cycle_matrix(self):
        """(Nf, Nj) face cycle matrix A (CSC, entries +-1)."""
        return self.cycle_matrix

    def get_cut_matrix(self):
        """(Nn, Nj) cut matrix M (CSC, entries +-1)."""
        return self.cut_matrix

    def _Mr(self):
        if self.cut_matrix_reduced is None:
            self.cut_matrix_reduced = self.cut_matrix[:-1, :]
        return self.cut_matrix_reduced

    def get_node_coordinates(self):
        return self.graph.coo()

    def get_junction_nodes(self):
        return self.graph.get_edges()

    def get_juncion_coordinates(self):
        x, y = self.get_node_coordinates()
        n1, n2 = self.get_junction_nodes()
        return x[n1], y[n1], x[n2], y[n2]

    get_junction_coordinates = get_juncion_coordinates

    def get_faces(self):
        return self.graph.get_face_cycles(to_list=True)[0]

    def get_face_areas(self):
        return self.graph.get_face_areas()

    def get_face_centroids(self):
        return self.graph.get_face_centroids()

    def locate_faces(self, x, y):
        """Index of the face whose centroid is nearest to each (x, y)."""
        if self.locator is None:
            self.locator = scipy.spatial.KDTree(np.stack(self.get_face_centroids(), axis=-1))
        _, faces = self.locator.query(np.stack(np.broadcast_arrays(x, y), axis=-1), k=1)
        return faces

    # ------------------------------------------------------------ solves
    def Msq_solve(self, b):
        """Solve M M^T x = b with the last node grounded (josephson_circuit.py:93-109)."""
        if self._Msq_factorized is None:
            Mr = self._Mr()
            self._Msq_factorized = scipy.sparse.linalg.factorized(Mr @ Mr.T)
        return np.append(self._Msq_factorized(b[:-1, ...]), np.zeros(b[-1:, ...].shape), axis=0)

    def Asq_solve(self, b):
        """Solve A A^T x = b (josephson_circuit.py:111-128)."""
        if self._Asq_factorized is None:
            A = self.cycle_matrix
            self._Asq_factorized = scipy.sparse.linalg.factorized(A @ A.T)
        return self._Asq_factorized(b)

    # ------------------------------------------------------------ editing
    def copy(self):
        n1, n2 = self.get_junction_nodes()
        return Circuit(EmbeddedGraph(self.graph.x, self.graph.y, n1, n2),
                       critical_current=self.critical_current, resistance=self.resistance,
                       capacitance=self.capacitance, inductance=self.inductance)

    def remove_nodes(self, nodes_to_remove):
        nodes_to_remove = np.asarray(nodes_to_remove).ravel()
        if nodes_to_remove.dtype == bool:
            mask = nodes_to_remove
        else:
            mask = np.zeros(self._Nn(), dtype=bool)
            mask[nodes_to_remove.astype(int)] = True
        n1, n2 = self.get_junction_nodes()
        keep = ~(mask[n1] | mask[n2])
        new_id = np.cumsum(~mask) - 1
        x, y = self.get_node_coordinates()
        L = self.inductance.tocsr()[keep, :][:, keep]
        return Circuit(EmbeddedGraph(x[~mask], y[~mask], new_id[n1[keep]], new_id[n2[keep]]),
                       critical_current=self.critical_current[keep],
                       resistance=self.resistance[keep], capacitance=self.capacitance[keep],
                       inductance=L)

    def remove_junctions(self, junctions_to_remove):
        j = np.asarray(junctions_to_remove).ravel()
        if j.dtype == bool:
            mask = j
        else:
            mask = np.zeros(self._Nj(), dtype=bool)
            mask[j.astype(int)] = True
        n1, n2 = self.get_junction_nodes()
        keep = ~mask
        L = self.inductance.tocsr()[keep, :][:, keep]
        return Circuit(EmbeddedGraph(self.graph.x, self.graph.y, n1[keep], n2[keep]),
                       critical_current=self.critical_current[keep],
                       resistance=self.resistance[keep], capacitance=self.capacitance[keep],
                       inductance=L)

    # ------------------------------------------------------------ save/load
    def save(self, filename):
        """Sequential ``np.save`` records, the reference's circuit file layout
        (josephson_circuit.py:643-663)."""
        with open(filename, "wb") as fh:
            self._save_records(fh)

    def _save_records(self, fh):
        x, y = self.graph.coo()
        n1, n2 = self.graph.get_edges()
        for arr in (x, y, n1, n2, self.critical_current, self.resistance, self.capacitance):
            np.save(fh, arr)
        L = self.inductance
        sparse = not getattr(self, "_inductance_dense", False)
        np.save(fh, sparse)
        if not sparse:
            L = np.ascontiguousarray(L.toarray())
        if sparse:
            L = scipy.sparse.csc_matrix(L)
            np.save(fh, L.indptr)
            np.save(fh, L.indices)
            np.save(fh, L.data)
        else:
            np.save(fh, L)

    @staticmethod
    def load(filename):
        """Read a file written by ``save`` (josephson_circuit.py:665-684)."""
        with open(filename, "rb") as fh:
            return Circuit._load_records(fh)

    @staticmethod
    def _load_records(fh):
        x, y, n1, n2, Ic, R, C = (np.load(fh) for _ in range(7))
        sparse = bool(np.load(fh))
        if sparse:
            indptr, indices, data = np.load(fh), np.load(fh), np.load(fh)
            Nj = len(n1)
            L = scipy.sparse.csc_matrix((data, indices, indptr), shape=(Nj, Nj))
        else:
            L = np.load(fh)
        return Circuit(EmbeddedGraph(x, y, n1, n2), critical_current=Ic, resistance=R,
                       capacitance=C, inductance=L)

    def _A_solve(self, b):
        """A x = b with x integral for integral b (josephson_circuit.py:782-801):
        the graph's integral cycle-space solve."""
        return self.graph._cycle_space_solve_for_integral_x(b)

    def _has_identical_critical_current(self):
        Ic = self.get_critical_current()
        return bool(np.allclose(Ic, Ic[0]))

    def _has_only_identical_self_inductance(self):
        if self._has_only_self_inductance():
            L = self.inductance.diagonal()
            return bool(np.allclose(L, L[0]))
        return False

    def __str__(self):
        return f"Circuit(Nn={self._Nn()}, Nj={self._Nj()}, Nf={self._Nf()})"


def _is_positive_definite_sparse(L):
    try:
        lu = scipy.sparse.linalg.splu(scipy.sparse.csc_matrix(L), diag_pivot_thresh=0.0,
                                      permc_spec="MMD_AT_PLUS_A",
                                      options={"SymmetricMode": True})
    except RuntimeError:
        return False
    return bool(np.all(lu.U.diagonal() > 0))


def node_to_junction_current(circuit: Circuit, node_current):
    """Junction currents whose node injections equal ``node_current``
    (static_problem.py:1174-1192): minimum-norm solution -M^T (M M^T)^-1 b."""
    return -circuit.get_cut_matrix().T @ circuit.Msq_solve(np.asarray(node_current))


class Lattice:
    """Mixin with the uniform-bias helper (josephson_circuit.py:887-944)."""

    def __init__(self, count_x, count_y, x_scale=1.0, y_scale=1.0):
        self.count_x = count_x
        self.count_y = count_y
        self.x_scale = x_scale
        self.y_scale = y_scale

    def get_count_x(self):
        return self.count_x

    def get_count_y(self):
        return self.count_y

    def get_x_scale(self):
        return self.x_scale

    def get_y_scale(self):
        return self.y_scale

    def current_base(self, angle, type="junction"):
        """Sources for a uniform current at ``angle``: node injections on opposite
        edges of the lattice, converted to junction sources by the minimum-norm
        cut-space solve when ``type == "junction"``."""
        if type == "junction":
            return node_to_junction_current(self, self.current_base(angle, type="node"))
        if type == "node":
            return self._I_node_h() * np.cos(angle) + self._I_node_v() * np.sin(angle)
        raise ValueError("type must be 'junction' or 'node'")


class SquareArray(Circuit, Lattice):
    def __init__(self, count_x, count_y, x_scale=1.0, y_scale=1.0):
        Lattice.__init__(self, count_x, count_y, x_scale, y_scale)
        Circuit.__init__(self, EmbeddedSquareGraph(count_x, count_y, x_scale, y_scale))

    def _I_node_h(self):
        x, _ = self.get_node_coordinates()
        return (x == 0).astype(int) - np.isclose(x, (self.count_x - 1) * self.x_scale).astype(int)

    def _I_node_v(self):
        _, y = self.get_node_coordinates()
        return (y == 0).astype(int) - np.isclose(y, (self.count_y - 1) * self.y_scale).astype(int)


class HoneycombArray(Circuit, Lattice):
    def __init__(self, count_x, count_y, x_scale=1.0, y_scale=1.0):
        Lattice.__init__(self, count_x, count_y, x_scale, y_scale)
        Circuit.__init__(self, EmbeddedHoneycombGraph(count_x, count_y, x_scale, y_scale))

    def _I_node_h(self):
        x, _ = self.get_node_coordinates()
        n = self.count_y / (self.count_y - 0.5)
        right = np.isclose(x, (3 * self.count_x - 1) * self.x_scale)
        return ((x == 0).astype(int) - right.astype(int)) * n

    def _I_node_v(self):
        _, y = self.get_node_coordinates()
        low = y < 0.1 * np.sqrt(3) * self.y_scale
        high = y > (self.count_y - 0.6) * np.sqrt(3) * self.y_scale
        return (low.astype(int) - high.astype(int)) * 1.5


class TriangularArray(Circuit, Lattice):
    def __init__(self, count_x, count_y, x_scale=1.0, y_scale=1.0):
        Lattice.__init__(self, count_x, count_y, x_scale, y_scale)
        Circuit.__init__(self, EmbeddedTriangularGraph(count_x, count_y, x_scale, y_scale))

    def _I_node_h(self):
        x, _ = self.get_node_coordinates()
        left = x < 0.1 * self.x_scale
        right = x > (self.count_x - 0.6) * self.x_scale
        return (left.astype(int) - right.astype(int)) * np.sqrt(3)

    def _I_node_v(self):
        _, y = self.get_node_coordinates()
        top = np.isclose(y, (self.count_y - 0.5) * np.sqrt(3) * self.y_scale)
        return (y == 0).astype(int) - top.astype(int)


class SQUID(Circuit):
    """Square of four junctions; vertical junctions have Ic=1000
    (josephson_circuit.py:1001-1020)."""

    def __init__(self):
        graph = EmbeddedGraph([0, 1, 1, 0], [0, 0, 1, 1], [0, 1, 2, 0], [1, 2, 3, 3])
        super().__init__(graph, critical_current=[1, 1000, 1, 1000])

    def horizontal_junctions(self):
        return np.array([1, 0, -1, 0])

    def vertical_junctions(self):
        return np.array([0, 1, 0, 1])
