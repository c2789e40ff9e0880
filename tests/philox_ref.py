"""numpy restatement of the SGLE noise generator the library uses:
Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC 2011) on the counter (i, 0, draw_lo, draw_hi) with key
(seed_lo, seed_hi), 53-bit uniforms u1 in (0, 1] and u2 in [0, 1) from the
two 64-bit words, and the Box-Muller transform.  Test helper."""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32) for x in (c0, c1, c2, c3))
    k0, k1 = np.uint32(k0), np.uint32(k1)
    for _ in range(10):
        p0 = M0 * c0.astype(np.uint64)
        p1 = M1 * c2.astype(np.uint64)
        hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & MASK).astype(np.uint32)
        hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & MASK).astype(np.uint32)
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
    return c0, c1, c2, c3


def sgle_normals(N, draw, seed):
    """complex standard normals v1 + i v2 at the N*N grid points (y, x)"""
    i = np.arange(N * N, dtype=np.uint64)
    c = philox4x32_10((i & MASK).astype(np.uint32), (i >> np.uint64(32)).astype(np.uint32),
                      np.full(N * N, draw & 0xFFFFFFFF, np.uint32), np.full(N * N, draw >> 32, np.uint32),
                      seed & 0xFFFFFFFF, seed >> 32)
    a = (c[1].astype(np.uint64) << np.uint64(32)) | c[0].astype(np.uint64)
    b = (c[3].astype(np.uint64) << np.uint64(32)) | c[2].astype(np.uint64)
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    return (r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)).reshape(N, N)
