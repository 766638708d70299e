"""Independent numpy restatement of the workload generators, used to pin the
C oracle's (and the HIP library's) generators bit-exactly on small inputs.

R-MAT: edge i descends `scale` quadrant levels; level l takes 32 random bits
from draw(i, l // 2) = mix64(key + (16 i + l // 2 + 1) * GOLDEN) (low half for
even l, high half for odd l) and picks quadrant a/b/c/d by integer thresholds
floor(p * 2^32).  mix64 is the SplitMix64 finaliser (Steele, Lea, Flood 2014).
Weights: lo + ((mix64(wkey + (u << 32 | v) * GOLDEN) >> 32) * span >> 32).
Grid: undirected lattice edges e (horizontal first), dropped when
mix64(gkey + (e + 1) * GOLDEN) >> 32 < floor(drop * 2^32).
"""
import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _key(seed, salt):
    return int(mix64(np.uint64((seed * GOLDEN + salt) & M64)))


def _thr(p):
    t = p * 4294967296.0
    return 0 if t <= 0 else (0xFFFFFFFF if t >= 4294967295.0 else int(t))


def rmat_edges(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1):
    m = (1 << scale) * edge_factor
    key = _key(seed, 0x5851F42D4C957F2D)
    tA, tAB, tABC = _thr(a), _thr(a + b), _thr(a + b + c)
    i = np.arange(m, dtype=np.uint64)
    src = np.zeros(m, np.uint64)
    dst = np.zeros(m, np.uint64)
    with np.errstate(over="ignore"):
        for l in range(scale):
            ctr = i * np.uint64(16) + np.uint64(l // 2 + 1)
            r = mix64(np.uint64(key) + ctr * np.uint64(GOLDEN))
            x = (r >> np.uint64(32 * (l & 1))) & np.uint64(0xFFFFFFFF)
            q = np.where(x < tA, 0, np.where(x < tAB, 1, np.where(x < tABC, 2, 3))).astype(np.uint64)
            src = (src << np.uint64(1)) | (q >> np.uint64(1))
            dst = (dst << np.uint64(1)) | (q & np.uint64(1))
    return src.astype(np.uint32), dst.astype(np.uint32)


def edge_weight(seed, u, v, lo, hi, symmetric=False):
    u = np.asarray(u, np.uint64)
    v = np.asarray(v, np.uint64)
    if symmetric:
        u, v = np.minimum(u, v), np.maximum(u, v)
    key = _key(seed, 0x2545F4914F6CDD1D)
    with np.errstate(over="ignore"):
        h = mix64(np.uint64(key) + ((u << np.uint64(32)) | v) * np.uint64(GOLDEN))
        span = np.uint64(hi - lo + 1)
        return (np.uint64(lo) + (((h >> np.uint64(32)) * span) >> np.uint64(32))).astype(np.uint32)


def grid_edges(n, drop=0.1, seed=1):
    key = _key(seed, 0x94D049BB133111EB)
    nh = n * (n - 1)
    e = np.arange(2 * nh, dtype=np.uint64)
    r_h, c_h = e[:nh] // np.uint64(n - 1), e[:nh] % np.uint64(n - 1)
    f = e[nh:] - np.uint64(nh)
    r_v, c_v = f // np.uint64(n), f % np.uint64(n)
    u = np.concatenate([r_h * np.uint64(n) + c_h, r_v * np.uint64(n) + c_v])
    v = np.concatenate([u[:nh] + np.uint64(1), u[nh:] + np.uint64(n)])
    with np.errstate(over="ignore"):
        x = mix64(np.uint64(key) + (e + np.uint64(1)) * np.uint64(GOLDEN)) >> np.uint64(32)
    keep = x >= np.uint64(_thr(drop))
    u, v = u[keep], v[keep]
    src = np.empty(2 * len(u), np.uint64)
    dst = np.empty(2 * len(u), np.uint64)
    src[0::2], dst[0::2] = u, v
    src[1::2], dst[1::2] = v, u
    return src.astype(np.uint32), dst.astype(np.uint32)


def csr_from_edges(V, src, dst, symmetrize=False):
    src = np.asarray(src, np.uint64)
    dst = np.asarray(dst, np.uint64)
    if symmetrize:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    keep = src != dst
    key = np.unique((src[keep] << np.uint64(32)) | dst[keep])
    s = (key >> np.uint64(32)).astype(np.int64)
    d = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, s + 1, 1)
    return np.cumsum(off).astype(np.uint64), d
