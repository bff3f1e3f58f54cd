"""Planar embedded graphs and their face (cycle) and cut matrices.

Host-side topology for Josephson circuits. The face order and the orientation of
the cycle matrix reproduce the reference exactly, because vortex configurations
(n = -A round(theta / 2pi)) are reported per face:

* edges are normalised to (low node, high node) and sorted lexicographically;
  sorted edge k traversed low->high is directed edge k, high->low is k + E
  (embedded_graph.py:861-867, 937-985);
* at every node the neighbours are ordered by angle; a walk that arrives at node v
  from u leaves towards the neighbour that precedes u in that angular order.
  The orbits of this map on the 2E directed edges are the l-cycles
  (embedded_graph.py:937-1033);
* cycles with non-positive signed area are the outer boundary and are dropped
  (embedded_graph.py:913-916);
* faces are ordered by cycle length, then by the smallest directed-edge id in the
  cycle, and each face's edge list starts at that edge (embedded_graph.py:1026-1032);
* A[f, j] = +1 if face f (counter-clockwise) passes junction j from its first to
  its second node as the user gave them, else -1 (embedded_graph.py:518-539);
* M[n, j] = -1 at the junction's first node and +1 at its second node
  (embedded_graph.py:541-554).

The implementation is vectorised numpy: cycle labels come from pointer doubling
on the successor permutation instead of the reference's step-by-step walk.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse
import scipy.sparse.csgraph

__all__ = ["EmbeddedGraph", "EmbeddedSquareGraph", "EmbeddedHoneycombGraph",
           "EmbeddedTriangularGraph", "NotSingleComponentError", "NotPlanarEmbeddingError",
           "NonSimpleError", "SelfLoopError", "NodeNotExistError", "EdgeNotExistError"]


class NotSingleComponentError(Exception):
    pass


class NotPlanarEmbeddingError(Exception):
    pass


class NonSimpleError(Exception):
    pass


class SelfLoopError(Exception):
    pass


class NodeNotExistError(Exception):
    pass


class EdgeNotExistError(Exception):
    pass


class EmbeddedGraph:
    """Graph with 2D node coordinates (embedded_graph.py:31-110).

    Parameters
    ----------
    x, y : (N,) arrays of node coordinates.
    node1, node2 : (E,) int arrays; endpoints of each edge (user orientation).
    require_single_component, require_planar_embedding : bool
        Raise if the graph is not connected / not a planar embedding.
    """

    def __init__(self, x, y, node1, node2, require_single_component=False,
                 require_planar_embedding=False):
        self.x = np.array(x, dtype=np.double).ravel()
        self.y = np.array(y, dtype=np.double).ravel()
        if self.x.shape != self.y.shape:
            raise ValueError("x and y must be same size")
        self.N = len(self.x)
        n1 = np.array(node1, dtype=np.int64).ravel()
        n2 = np.array(node2, dtype=np.int64).ravel()
        if n1.shape != n2.shape:
            raise ValueError("node1 and node2 must be same size")
        self.E = len(n1)
        if self.E and (min(n1.min(), n2.min()) < 0 or max(n1.max(), n2.max()) >= self.N):
            raise NodeNotExistError("edge refers to a node that does not exist")
        if np.any(n1 == n2):
            raise SelfLoopError("no edge is allowed to have identical end nodes.")
        self._user_n1, self._user_n2 = n1, n2
        # normalised orientation and lexicographic order of the edges
        self._flip = n1 > n2
        lo = np.minimum(n1, n2)
        hi = np.maximum(n1, n2)
        self._order = np.lexsort((hi, lo))          # sorted position -> user edge
        self._rank = np.empty(self.E, dtype=np.int64)  # user edge -> sorted position
        self._rank[self._order] = np.arange(self.E)
        self._lo, self._hi = lo[self._order], hi[self._order]
        if self.E > 1 and np.any((self._lo[1:] == self._lo[:-1]) & (self._hi[1:] == self._hi[:-1])):
            raise NonSimpleError("no duplicate edges allowed.")
        self._faces = None
        if require_single_component:
            self._assert_single_component()
        if require_planar_embedding:
            self._assert_planar_embedding()

    # ------------------------------------------------------------------ basics
    def coo(self):
        return self.x, self.y

    def node_count(self):
        return self.N

    def edge_count(self):
        return self.E

    def get_edges(self):
        """Endpoints of all edges in the user's order and orientation."""
        return self._user_n1.copy(), self._user_n2.copy()

    def adjacency_matrix(self):
        n1, n2 = self._user_n1, self._user_n2
        data = np.ones(2 * self.E)
        return scipy.sparse.coo_matrix((data, (np.append(n1, n2), np.append(n2, n1))),
                                       shape=(self.N, self.N)).tocsr()

    def get_num_components(self):
        return scipy.sparse.csgraph.connected_components(self.adjacency_matrix(), directed=False,
                                                         return_labels=False)

    def get_components(self):
        return scipy.sparse.csgraph.connected_components(self.adjacency_matrix(), directed=False,
                                                         return_labels=True)[1]

    # ------------------------------------------------------------ face search
    def _successor(self):
        """Successor of every directed edge in the face walk, shape (2E,)."""
        E = self.E
        # directed edge d: tail = start node, head = end node
        tail = np.concatenate((self._lo, self._hi))
        head = np.concatenate((self._hi, self._lo))
        ang = np.arctan2(self.y[head] - self.y[tail], self.x[head] - self.x[tail])
        # outgoing directed edges grouped by tail node, ascending angle
        out_sorted = np.lexsort((ang, tail))
        counts = np.bincount(tail, minlength=self.N)
        start = np.concatenate(([0], np.cumsum(counts)[:-1]))
        slot = np.empty(2 * E, dtype=np.int64)      # position of each directed edge in its node ring
        slot[out_sorted] = np.arange(2 * E)
        # arriving along d (tail u -> head v), the reverse edge v->u sits in v's ring;
        # leave along the edge one position earlier in that ring (cyclically).
        rev = np.concatenate((np.arange(E, 2 * E), np.arange(E)))
        r = rev
        v = head
        pos = slot[r] - start[v]
        prev = (pos - 1) % counts[v]
        return out_sorted[start[v] + prev]

    def _compute_faces(self):
        if self._faces is not None:
            return self._faces
        E = self.E
        if E == 0:
            self._faces = (np.zeros(1, dtype=np.int64), np.zeros(0, dtype=np.int64),
                           np.zeros(0), np.zeros(0, dtype=np.int64))
            return self._faces
        nxt = self._successor()
        # cycle label = smallest directed id on the cycle (pointer doubling)
        lab = np.arange(2 * E)
        jump = nxt.copy()
        while True:
            new = np.minimum(lab, lab[jump])
            if np.array_equal(new, lab):
                break
            lab = new
            jump = jump[jump]
        self._lab = lab
        labels, inverse, lengths = np.unique(lab, return_inverse=True, return_counts=True)
        tail = np.concatenate((self._lo, self._hi))
        head = np.concatenate((self._hi, self._lo))
        det = self.x[tail] * self.y[head] - self.x[head] * self.y[tail]
        area = 0.5 * np.bincount(inverse, weights=det, minlength=len(labels))
        keep = (area > 0) & ~np.isclose(area, 0.0)
        order = np.lexsort((labels[keep], lengths[keep]))
        f_label = labels[keep][order]
        f_len = lengths[keep][order]
        f_area = area[keep][order]
        # walk every kept cycle from its label
        F = len(f_label)
        ptr = np.concatenate(([0], np.cumsum(f_len)))
        edges = np.empty(int(ptr[-1]), dtype=np.int64)
        cur = f_label.copy()
        active = np.arange(F)
        step = 0
        while len(active):
            edges[ptr[active] + step] = cur
            step += 1
            alive = f_len[active] > step
            active = active[alive]
            cur = nxt[cur[alive]]
        l_count = len(labels)
        self._faces = (ptr, edges, f_area, l_count)
        return self._faces

    def face_count(self):
        return len(self._compute_faces()[0]) - 1

    # ------------------------------------------------- l-cycles (with boundary)
    def _l_cycles(self):
        """Every cycle of the face walk, boundary cycles included, in the
        reference's l-cycle order (embedded_graph.py:987-1033: length, then the
        smallest directed edge id on the cycle). Returns (l-cycle of every
        directed edge, the cycles' directed edges in walk order from their
        smallest id as (ptr, edges), indices of the boundary cycles: area not
        positive, embedded_graph.py:913-916)."""
        if getattr(self, "_lcyc", None) is not None:
            return self._lcyc
        self._compute_faces()
        E = self.E
        nxt = self._successor()
        labels, inverse, lengths = np.unique(self._lab, return_inverse=True, return_counts=True)
        order = np.lexsort((labels, lengths))
        rank = np.empty(len(labels), dtype=np.int64)
        rank[order] = np.arange(len(labels))
        cyc_of_dir = rank[inverse]
        tail = np.concatenate((self._lo, self._hi))
        head = np.concatenate((self._hi, self._lo))
        det = self.x[tail] * self.y[head] - self.x[head] * self.y[tail]
        area = 0.5 * np.bincount(cyc_of_dir, weights=det, minlength=len(labels))
        boundary = np.flatnonzero((area < 0) | np.isclose(area, 0.0))
        f_len = lengths[order]
        ptr = np.concatenate(([0], np.cumsum(f_len)))
        edges = np.empty(2 * E, dtype=np.int64)
        cur = labels[order].copy()
        active = np.arange(len(order))
        step = 0
        while len(active):
            edges[ptr[active] + step] = cur
            step += 1
            alive = f_len[active] > step
            active = active[alive]
            cur = nxt[cur[alive]]
        self._lcyc = (cyc_of_dir, ptr, edges, boundary)
        return self._lcyc

    def l_cycle_matrix(self):
        """(L, E) cycle matrix of all l-cycles, boundary included, one entry per
        traversed edge in walk order (embedded_graph.py:505-523)."""
        _, ptr, edges, _ = self._l_cycles()
        E = self.E
        cols = self._order[edges % E]
        data = np.where(edges < E, 1, -1)
        data = np.where(self._flip[cols], -data, data)
        return scipy.sparse.csr_matrix((data, cols, ptr), shape=(len(ptr) - 1, E)).tocsc()

    def get_common_edge_of_l_cycles(self, cycle1, cycle2, return_orientation=False):
        """An edge both l-cycles contain, the lowest in sorted edge order, or -1
        (embedded_graph.py:569-612); with return_orientation, whether cycle1
        passes it counter-clockwise from its lowest-index node."""
        cyc_of_dir = self._l_cycles()[0]
        E = self.E
        f1, f2 = cyc_of_dir[:E], cyc_of_dir[E:2 * E]
        mask = f1 < f2
        lo, hi = np.where(mask, f1, f2), np.where(mask, f2, f1)
        sorter = np.lexsort((hi, lo))
        key = lo[sorter] * (len(cyc_of_dir) + 1) + hi[sorter]
        c1 = np.asarray(cycle1, dtype=np.int64)
        c2 = np.asarray(cycle2, dtype=np.int64)
        shape = c1.shape
        c1, c2 = c1.ravel(), c2.ravel()
        mask2 = c1 < c2
        fmin, fmax = np.where(mask2, c1, c2), np.where(mask2, c2, c1)
        pos = np.searchsorted(key, fmin * (len(cyc_of_dir) + 1) + fmax)
        pos[pos >= E] = E - 1
        found = key[pos] == fmin * (len(cyc_of_dir) + 1) + fmax
        orient = mask[sorter][pos] ^ ~mask2
        ids = self._order[sorter[pos]]
        ids = np.where(found, ids, -1).reshape(shape)
        if return_orientation:
            orient = np.where(found, orient, False).reshape(shape)
            return ids, orient ^ np.where(ids >= 0, self._flip[np.maximum(ids, 0)], False)
        return ids

    def _cycle_space_solve_for_integral_x(self, b):
        """A x = b (A: the face cycle matrix) with x integral when b is
        (embedded_graph.py:1041-1090): a depth-first tree of the dual graph from
        the boundary cycle; each face's tally passes to its parent across their
        common edge, leaves first."""
        if self.get_num_components() != 1 or not self.is_planar_embedding():
            raise ValueError("only implemented for single component planar embedding")
        _, ptr, _, boundary = self._l_cycles()
        E, F = self.E, self.face_count()
        b = np.array(b, dtype=np.double)
        shape = list(b.shape)
        b = b.reshape(-1, F)
        b_idx = int(boundary[0])
        tally = np.concatenate((b[:, :b_idx], np.zeros((b.shape[0], 1)), b[:, b_idx:]), axis=1)
        Al = self.l_cycle_matrix()
        cur, pred = scipy.sparse.csgraph.depth_first_order(Al @ Al.T, b_idx)
        prev = pred[cur]
        prev[0] = -1
        juncs, orient = self.get_common_edge_of_l_cycles(cur, prev, return_orientation=True)
        sgn = -1 + 2 * orient.astype(np.double)
        x = np.zeros((b.shape[0], E))
        for i in reversed(range(F + 1)):
            if prev[i] >= 0:
                tally[:, prev[i]] += tally[:, cur[i]]
                x[:, juncs[i]] += tally[:, cur[i]] * sgn[i]
                tally[:, cur[i]] = 0
        keep = np.ones(len(ptr) - 1, dtype=bool)
        keep[boundary] = False
        if not np.allclose((Al @ x.T)[keep, :], b.T):
            raise ValueError("failed integral solve of cycle space linear problem")
        shape[-1] = E
        return x.reshape(tuple(shape))

    def l_cycle_count(self):
        return int(self._compute_faces()[3])

    def is_planar_embedding(self):
        return self.node_count() + self.face_count() == self.edge_count() + 1

    def _assert_single_component(self):
        if self.get_num_components() != 1:
            raise NotSingleComponentError()

    def _assert_planar_embedding(self):
        if not self.is_planar_embedding():
            raise NotPlanarEmbeddingError()

    def _face_tail_nodes(self):
        ptr, edges, _, _ = self._compute_faces()
        E = self.E
        k = edges % E
        return ptr, np.where(edges < E, self._lo[k], self._hi[k])

    def get_face_cycles(self, to_list=True):
        """Nodes met walking each face counter-clockwise (embedded_graph.py:223-247)."""
        ptr, nodes = self._face_tail_nodes()
        lengths = np.diff(ptr)
        if to_list:
            return [nodes[ptr[i]:ptr[i + 1]].tolist() for i in range(len(lengths))], lengths
        return nodes, lengths

    def get_face_areas(self):
        return self._compute_faces()[2].copy()

    def get_face_centroids(self):
        """Area centroids of the faces (polygon formula)."""
        ptr, nodes = self._face_tail_nodes()
        F = len(ptr) - 1
        lengths = np.diff(ptr)
        fid = np.repeat(np.arange(F), lengths)
        nxt_pos = np.arange(len(nodes)) + 1
        nxt_pos[ptr[1:] - 1] = ptr[:-1]
        a, b = nodes, nodes[nxt_pos]
        det = self.x[a] * self.y[b] - self.x[b] * self.y[a]
        area6 = 3.0 * np.bincount(fid, weights=det, minlength=F)
        cx = np.bincount(fid, weights=(self.x[a] + self.x[b]) * det, minlength=F) / area6
        cy = np.bincount(fid, weights=(self.y[a] + self.y[b]) * det, minlength=F) / area6
        return cx, cy

    def face_cycle_matrix(self):
        """Sparse (F, E) cycle matrix A in CSC format (embedded_graph.py:525-539).

        Entries are kept per traversed edge (a junction walked in both directions by
        one face appears twice, with opposite signs), as the reference builds it.
        """
        ptr, edges, _, _ = self._compute_faces()
        E = self.E
        cols = self._order[edges % E]
        data = np.where(edges < E, 1, -1)
        data = np.where(self._flip[cols], -data, data)
        return scipy.sparse.csr_matrix((data, cols, ptr), shape=(len(ptr) - 1, E)).tocsc()

    def cut_space_matrix(self):
        """Sparse (N, E) cut matrix M in CSC format (embedded_graph.py:541-554)."""
        E, N = self.E, self.N
        n1, n2 = self._user_n1, self._user_n2
        row = np.concatenate((n1, n2))
        col = np.concatenate((np.arange(E), np.arange(E)))
        data = np.concatenate((-np.ones(E), np.ones(E)))
        return scipy.sparse.coo_matrix((data, (row, col)), shape=(N, E)).tocsc()

    # -------------------------------------------------------------- editing
    def remove_nodes(self, node_ids):
        mask = np.zeros(self.N, dtype=bool)
        mask[np.asarray(node_ids)] = True
        keep_e = ~(mask[self._user_n1] | mask[self._user_n2])
        new_id = np.cumsum(~mask) - 1
        return EmbeddedGraph(self.x[~mask], self.y[~mask], new_id[self._user_n1[keep_e]],
                             new_id[self._user_n2[keep_e]])


class EmbeddedSquareGraph(EmbeddedGraph):
    """Square lattice of count_x by count_y nodes (embedded_graph.py:1093-1111).

    Node (ix, iy) has index iy*count_x + ix. Junctions: all horizontal edges in
    row-major order, then all vertical edges in row-major order.
    """

    def __init__(self, count_x, count_y, x_scale=1.0, y_scale=1.0):
        iy, ix = np.divmod(np.arange(count_x * count_y), count_x)
        grid = np.arange(count_x * count_y).reshape(count_y, count_x)
        n1 = np.concatenate((grid[:, :-1].ravel(), grid[:-1, :].ravel()))
        n2 = np.concatenate((grid[:, 1:].ravel(), grid[1:, :].ravel()))
        super().__init__(ix * x_scale, iy * y_scale, n1, n2)


class EmbeddedHoneycombGraph(EmbeddedGraph):
    """Honeycomb lattice built from a 4-node unit cell (embedded_graph.py:1114-1140).

    Cell (cx, cy) holds nodes at (3cx, sqrt3 cy), (3cx+0.5, sqrt3 cy+sqrt.75),
    (3cx+1.5, sqrt3 cy+sqrt.75), (3cx+2, sqrt3 cy); node blocks are stored
    one sub-node type after another.
    """

    def __init__(self, count_x, count_y, x_scale=1.0, y_scale=1.0):
        cy, cx = np.divmod(np.arange(count_x * count_y), count_x)
        bx, by = 3.0 * cx, np.sqrt(3.0) * cy
        h = np.sqrt(0.75)
        xs = np.concatenate((bx, bx + 0.5, bx + 1.5, bx + 2.0))
        ys = np.concatenate((by, by + h, by + h, by))
        g = np.arange(count_x * count_y).reshape(count_y, count_x)
        s = count_x * count_y
        pairs = [(g, g + s), (g[:-1, :] + s, g[1:, :]), (g + s, g + 2 * s),
                 (g[:-1, :] + 2 * s, g[1:, :] + 3 * s), (g + 2 * s, g + 3 * s),
                 (g[:, :-1] + 3 * s, g[:, 1:])]
        n1 = np.concatenate([a.ravel() for a, _ in pairs])
        n2 = np.concatenate([b.ravel() for _, b in pairs])
        super().__init__(xs * x_scale, ys * y_scale, n1, n2)


class EmbeddedTriangularGraph(EmbeddedGraph):
    """Triangular lattice built from a 2-node unit cell (embedded_graph.py:1143-1166)."""

    def __init__(self, count_x, count_y, x_scale=1.0, y_scale=1.0):
        cy, cx = np.divmod(np.arange(count_x * count_y), count_x)
        bx, by = cx.astype(np.double), np.sqrt(3.0) * cy
        xs = np.concatenate((bx, bx + 0.5))
        ys = np.concatenate((by, by + np.sqrt(0.75)))
        g = np.arange(count_x * count_y).reshape(count_y, count_x)
        s = count_x * count_y
        pairs = [(g, g + s), (g[:, :-1], g[:, 1:]), (g[:-1, :] + s, g[1:, :]),
                 (g[:-1, :-1] + s, g[1:, 1:]), (g[:, :-1] + s, g[:, 1:] + s),
                 (g[:, 1:], g[:, :-1] + s)]
        n1 = np.concatenate([a.ravel() for a, _ in pairs])
        n2 = np.concatenate([b.ravel() for _, b in pairs])
        super().__init__(xs * x_scale, ys * y_scale, n1, n2)
