"""Direct solve of the face system of a uniform rectangular lattice.

The reference factorises the face matrix K = A (L + diag(1/Cnext)) A^T once
with SuperLU and applies the factor every step (time_evolution.py:476, 505-506,
560-569). For a rectangular lattice of R x C faces (SquareArray) whose
junctions have one weight per direction (identical R and C per junction, no
mutual inductance), K is a Kronecker sum. With face (r, c) at position
p = r C + c,

    K = g_r (T_R (x) I_C) + g_c (I_R (x) T_C),    T_n = tridiag(-1, 2, -1),

where g_r couples vertically adjacent faces (p, p + C) and g_c horizontally
adjacent ones (p, p + 1). T_C = S diag(mu) S with the orthonormal, symmetric
sine matrix (DST-I)

    S[i][j] = sqrt(2 / (C + 1)) sin(pi (i + 1)(j + 1) / (C + 1)),
    mu_k = 4 sin^2(pi (k + 1) / (2 (C + 1))),

so K^-1 = (I (x) S) M^-1 (I (x) S), M block diagonal in the sine mode k with
the symmetric positive definite tridiagonal block g_r T_R + g_c mu_k I along r.
That is Hockney's method (J. ACM 12 (1965) 95-113; Buzbee, Golub and Nielson,
SIAM J. Numer. Anal. 7 (1970) 627-656): a sine transform of every row, one
tridiagonal solve per mode, and the inverse transform. On the device the two
transforms are dense FP64 products of the rows with S; S[C-1-i][j] =
(-1)^j S[i][j] halves them: out[j] = sum_{i<H} F_j[i] S[i][j] with
F_j[i] = v[i] + (-1)^j v[C-1-i] (H = ceil(C/2); for odd C the middle index
pairs with itself, F = v there for even j, and S is 0 there for odd j). The
tridiagonal blocks are factorised once here (Thomas algorithm) and applied per
step: forward d_r = (y_r + g_r d_{r-1}) inv_r, backward x_r = d_r - up_r x_{r+1}
with up_r = (-g_r) inv_r formed on the fly (the same product the factorisation
forms), so only the pivots are tabulated. Both sweeps are linear recurrences
with the multiplier a_r = g_r inv_r; the device cuts each mode's rows into
blocks that run from a zero start, chains the block ends, and adds each
block's running product of multipliers times the carried value.

The solve is exact up to rounding, like the reference's SuperLU solve, and is
used only when the circuit's K equals the Kronecker form to within 8 ulp of its
largest entry (checked entry by entry).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse

from . import _native

# sine modes one wave accumulates at once (even, so a mode's parity is that of
# its slot); the sine table's rows are padded to a multiple of it
CHUNK = 14
# each mode's tridiagonal recurrence runs in min(R, TRI_THREADS * parts // C) blocks (column-wide)
TRI_THREADS = 1024
# dynamic LDS a split workgroup may take for its part-local tridiagonal solve
SPLIT_LDS_MAX = 163840 - 256


def _kron_form(R, C, g_r, g_c):
    I_R, I_C = scipy.sparse.identity(R), scipy.sparse.identity(C)
    T = lambda n: scipy.sparse.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    return (g_r * scipy.sparse.kron(T(R), I_C) + g_c * scipy.sparse.kron(I_R, T(C))).tocsr()


def lattice_form(K, xy):
    """(R, C, pos_of_face, g_r, g_c) when the face matrix K (Nf x Nf) of faces
    with centroids xy (Nf, 2) is the Kronecker form above on a rectangular grid
    of the centroids; None otherwise. The sine transform runs along the shorter
    side (C <= R)."""
    Nf = K.shape[0]
    if Nf < 4 or xy.shape != (Nf, 2):
        return None
    scale = max(np.abs(xy).max(), 1.0)
    xq = np.round(xy[:, 0] / scale, 9)
    yq = np.round(xy[:, 1] / scale, 9)
    xs, xi = np.unique(xq, return_inverse=True)
    ys, yi = np.unique(yq, return_inverse=True)
    nx, ny = len(xs), len(ys)
    if nx * ny != Nf or nx < 2 or ny < 2:
        return None
    if nx <= ny:
        R, C, pos = ny, nx, yi * nx + xi
    else:
        R, C, pos = nx, ny, xi * ny + yi
    if len(np.unique(pos)) != Nf:
        return None
    face_of_pos = np.empty(Nf, dtype=np.int64)
    face_of_pos[pos] = np.arange(Nf)
    Kg = scipy.sparse.csr_matrix(K)[face_of_pos][:, face_of_pos].tocsr()
    g_c = -Kg[0, 1]
    g_r = -Kg[0, C]
    if not (np.isfinite(g_r) and np.isfinite(g_c)) or g_r <= 0 or g_c <= 0:
        return None
    diff = (Kg - _kron_form(R, C, g_r, g_c)).tocsr()
    diff.eliminate_zeros()
    tol = 8 * np.finfo(np.float64).eps * np.abs(Kg.data).max()
    if diff.nnz and np.abs(diff.data).max() > tol:
        return None
    return R, C, pos.astype(np.int64), float(g_r), float(g_c)


class LatticeSolve:
    """Tables of the lattice solve of one circuit, in the layout the kernel
    reads: sine[ES][CP] (S[i][j] for i < H, zero padding) and inv[R][CP] (the
    inverse Thomas pivots of mode k's tridiagonal block at row r)."""

    def __init__(self, R, C, pos_of_face, g_r, g_c):
        self.R, self.C = int(R), int(C)
        self.g_r, self.g_c = g_r, g_c
        self.pos_of_face = np.asarray(pos_of_face, dtype=np.int64)
        self.face_of_pos = np.empty_like(self.pos_of_face)
        self.face_of_pos[self.pos_of_face] = np.arange(len(self.pos_of_face))
        H = (C + 1) // 2
        self.H = H
        self.ES = H + (H & 1)
        self.CP = -(-C // CHUNK) * CHUNK
        i = np.arange(self.ES)[:, None]
        j = np.arange(self.CP)[None, :]
        m = ((i + 1) * (j + 1)) % (2 * (C + 1))
        sine = np.sqrt(2.0 / (C + 1)) * np.sin(np.pi * m / (C + 1))
        sine[m % (C + 1) == 0] = 0.0                  # sin of a multiple of pi
        sine[(i >= H) | (j >= C)] = 0.0
        self.sine = np.ascontiguousarray(sine)
        k = np.arange(C)
        mu = 4.0 * np.sin(np.pi * (k + 1) / (2.0 * (C + 1))) ** 2
        diag = 2.0 * g_r + g_c * mu
        inv = np.zeros((R, self.CP))
        cprev = np.zeros(C)
        for r in range(R):
            # pivot_r = diag - off c'_{r-1}, off = -g_r, c'_{r-1} = off / pivot_{r-1} (c'_{-1} = 0)
            piv = diag - (-g_r) * cprev
            inv[r, :C] = 1.0 / piv
            cprev = (-g_r) * inv[r, :C]
        self.inv = np.ascontiguousarray(inv)

    def scratch_rows(self, Nf, parts=1, global_store=False):
        """Doubles after the face vector, at an even offset (16-byte loads): the
        folded rows (E and O, R x EF each; EF = ES in LDS, ES rounded up to 4 in
        a global store, which keeps them in blocks of four), which the
        tridiagonal solves reuse for their forward and backward block ends
        (values and products, 4 B C)."""
        EF = (self.ES + 3) & ~3 if global_store else self.ES
        return (Nf & 1) + max(2 * self.R * EF, 4 * self.column_tri_blocks(parts) * self.C)

    def column_tri_blocks(self, parts=1):
        """Blocks per mode of the column-wide tridiagonal solve (block ends in
        the face store): JJA_LATTICE_TRI_THREADS per part, over C."""
        return max(1, min(self.R, TRI_THREADS * parts // self.C))

    def part_local(self, parts):
        """(mode slots per part, blocks per mode) of the part-local tridiagonal
        solve a split column runs in each workgroup's LDS (all R rows of
        ceil(H / parts) mode pairs (i, C-1-i), two slots each, and 4 block-end
        arrays; holding both modes of a pair, the part also writes the second
        fold), or None when it does not fit (the column-wide solve then runs)."""
        if parts <= 1:
            return None
        mp = 2 * -(-self.H // parts)
        bp = max(1, min(self.R, TRI_THREADS // mp))
        need = 8 * (self.R * mp + 4 * bp * mp)
        return (mp, bp) if need <= SPLIT_LDS_MAX else None

    def split_lds_bytes(self, parts):
        """Dynamic LDS per workgroup of a split column's part-local solve (0 if none)."""
        pl = self.part_local(parts)
        return 0 if pl is None else 8 * (self.R * pl[0] + 4 * pl[1] * pl[0])

    def tri_blocks(self, parts=1):
        """Blocks each mode's tridiagonal recurrence is cut into, for a column
        worked by ``parts`` workgroups (the part-local solve's where it runs)."""
        pl = self.part_local(parts)
        return pl[1] if pl is not None else self.column_tri_blocks(parts)

    def table_bytes(self):
        """Bytes of the solve's tables (the tabulated half of the sine matrix
        and the pivots), which every column of a step reads."""
        return 8 * (self.H * self.C + self.R * self.C)

    def set_desc(self, d):
        P = _native.ptr
        d.lattice_rows, d.lattice_cols = self.R, self.C
        d.lattice_sine = P(self.sine, _native.C.c_double)
        d.lattice_inv = P(self.inv, _native.C.c_double)
        d.lattice_g = self.g_r

    # ------------------------------------------------------ host emulation
    def _fold(self, Z):
        C, H, ES = self.C, self.H, self.ES
        E = np.zeros((Z.shape[0], ES))
        O = np.zeros((Z.shape[0], ES))
        rev = Z[:, ::-1]
        E[:, :H] = Z[:, :H] + rev[:, :H]
        O[:, :H] = Z[:, :H] - rev[:, :H]
        if C % 2:
            E[:, H - 1] = Z[:, H - 1]
            O[:, H - 1] = 0.0
        return E, O

    def _transform(self, Z):
        E, O = self._fold(Z)
        out = np.empty_like(Z)
        out[:, 0::2] = E @ self.sine[:, :self.C][:, 0::2]
        out[:, 1::2] = O @ self.sine[:, :self.C][:, 1::2]
        return out

    def _tridiag(self, Y, parts=1):
        """Every mode's tridiagonal solve in place in Y (R x C), by the device's
        blocked recurrences."""
        R, C, g = self.R, self.C, self.g_r
        inv = self.inv[:, :C]
        B = self.tri_blocks(parts)
        m = -(-R // B)
        blocks = [(b * m, min(R, (b + 1) * m)) for b in range(B) if b * m < R]
        ends = []
        for r0, r1 in blocks:                       # forward, each block from zero
            d, pr = np.zeros(C), np.ones(C)
            for r in range(r0, r1):
                d = (Y[r] + g * d) * inv[r]
                pr = pr * (g * inv[r])
                Y[r] = d
            ends.append((d, pr))
        cin = np.zeros(C)
        for i, (r0, r1) in enumerate(blocks[1:], start=1):
            cin = ends[i - 1][1] * cin + ends[i - 1][0]
            pr = np.ones(C)
            for r in range(r0, r1):
                pr = pr * (g * inv[r])
                Y[r] = pr * cin + Y[r]
        ends = []
        for r0, r1 in blocks:                       # backward, each block from zero
            x, pr = np.zeros(C), np.ones(C)
            for r in range(r1 - 1, r0 - 1, -1):
                a = g * inv[r]
                x = a * x + Y[r]
                pr = pr * a
                Y[r] = x
            ends.append((x, pr))
        cin = np.zeros(C)
        for i in range(len(blocks) - 2, -1, -1):
            cin = ends[i + 1][1] * cin + ends[i + 1][0]
            r0, r1 = blocks[i]
            pr = np.ones(C)
            for r in range(r1 - 1, r0 - 1, -1):
                pr = pr * (g * inv[r])
                Y[r] = pr * cin + Y[r]

    def emulate(self, b, parts=1):
        """K^-1 b for b (Nf,) or (Nf, m) in circuit face order, by the kernel's
        steps for a column of ``parts`` workgroups (its sums agree with these
        to rounding)."""
        b = np.asarray(b, dtype=np.float64)
        vec = b.ndim == 1
        B = b[:, None] if vec else b
        R, C = self.R, self.C
        out = np.empty_like(B)
        for col in range(B.shape[1]):
            Z = B[self.face_of_pos, col].reshape(R, C)
            Y = self._transform(Z)
            self._tridiag(Y, parts)
            X = self._transform(Y)
            out[self.face_of_pos, col] = X.ravel()
        return out[:, 0] if vec else out


def lattice_solve(K, xy):
    """A LatticeSolve for the face matrix K of faces at centroids xy, or None
    when K is not the Kronecker form of a rectangular lattice."""
    form = lattice_form(K, xy)
    return None if form is None else LatticeSolve(*form)
