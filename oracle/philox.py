"""Counter-based Gaussian noise stream shared by the GPU step kernel and the oracle.

TEST INFRASTRUCTURE: only tests, the smoke check and the CPU baseline import this.

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11): a 4x32-bit counter and a 2x32-bit key go through 10 rounds of
  (hi0, lo0) = mulhilo(0xD2511F53, c0); (hi1, lo1) = mulhilo(0xCD9E8D57, c2)
  c <- (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0)
with the key bumped by (0x9E3779B9, 0xBB67AE85) between rounds.
Pinned by the published known-answer vectors in tests/test_oracle.py.

Normal deviate for junction j, column w, step i, seed s:
  counter = (j >> 1, w, i, 0), key = (s mod 2^32, s >> 32)
  u1 = ((r1:r0 as 64 bit) >> 11 + 1) * 2^-53   in (0, 1]
  u2 = ((r3:r2 as 64 bit) >> 11) * 2^-53       in [0, 1)
  rho = sqrt(-2 log u1), a = (2 pi) * u2
  Z = rho cos a   for even j,  rho sin a   for odd j     (Box-Muller)
"""
import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint64(0x9E3779B9)
_W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)
_S32 = np.uint64(32)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10. All inputs are uint32-valued arrays (any int dtype);
    returns four uint64 arrays holding 32-bit words."""
    c0, c1, c2, c3 = (np.asarray(v).astype(np.uint64) & _MASK for v in (c0, c1, c2, c3))
    k0 = np.asarray(k0).astype(np.uint64) & _MASK
    k1 = np.asarray(k1).astype(np.uint64) & _MASK
    for r in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> _S32, p0 & _MASK
        hi1, lo1 = p1 >> _S32, p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _MASK, lo1, (hi0 ^ c3 ^ k1) & _MASK, lo0
        if r < 9:
            k0 = (k0 + _W0) & _MASK
            k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def normals(seed, step, columns, n_junctions):
    """(n_junctions, len(columns)) standard normal deviates for one time step."""
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    j = np.arange(n_junctions, dtype=np.uint64)[:, None]
    w = np.asarray(columns, dtype=np.uint64)[None, :]
    r0, r1, r2, r3 = philox4x32_10(j >> np.uint64(1), w, np.uint64(step), np.uint64(0), k0, k1)
    a = (r0 | (r1 << _S32)) >> np.uint64(11)
    b = (r2 | (r3 << _S32)) >> np.uint64(11)
    u1 = (a + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = b.astype(np.float64) * 2.0 ** -53
    rho = np.sqrt(-2.0 * np.log(u1))
    ang = (2.0 * np.pi) * u2
    even = (np.arange(n_junctions) % 2 == 0)[:, None]
    return np.where(even, rho * np.cos(ang), rho * np.sin(ang))
