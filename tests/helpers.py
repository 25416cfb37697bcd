"""Instance generators and small utilities shared by the tests and bench.py.

Seeded generators (splitmix64 words) so that GPU and CPU paths see identical
couplings: +-1 from the top bit, integers in [lo, hi] by multiply-shift,
Gaussians by Box-Muller in FP64 rounded to FP32 and scaled by 1/sqrt(N).
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    x = seed & _M64
    for k in range(count):
        x = (x + 0x9E3779B97F4A7C15) & _M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out[k] = z ^ (z >> 31)
    return out


def splitmix64_np(seed: int, count: int) -> np.ndarray:
    """Vectorised splitmix64 (same stream as splitmix64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def ext_from_upper(J_upper: np.ndarray, h: np.ndarray) -> np.ndarray:
    """(n+1)^2 layout: J upper triangle, h in column n, 1 in the corner."""
    n = len(h)
    E = np.zeros((n + 1, n + 1))
    E[:n, :n] = np.triu(J_upper, 1)
    E[:n, n] = h
    E[n, n] = 1.0
    return E


def symmetric_ext(E: np.ndarray) -> np.ndarray:
    return np.triu(E) + np.triu(E, 1).T


def sk_pm1(n: int, seed: int) -> np.ndarray:
    """Dense SK with J_ij = +-1 (top bit), h = 0, as (n+1)^2 ext layout."""
    w = splitmix64_np(seed, n * n).reshape(n, n)
    J = np.where((w >> np.uint64(63)) == 1, 1.0, -1.0)
    return ext_from_upper(np.triu(J, 1), np.zeros(n))


def sk_gauss(n: int, seed: int, field: bool = False) -> np.ndarray:
    """Dense SK with J_ij ~ N(0, 1/n) rounded to FP32 (Box-Muller in FP64)."""
    w = splitmix64_np(seed, n * n + n + 2)
    u1 = ((w[0::2] >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = ((w[1::2] >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    g = np.concatenate([g, np.sqrt(-2.0 * np.log(u1)) * np.sin(2.0 * np.pi * u2)])
    J = g[: n * n].reshape(n, n).astype(np.float32).astype(np.float64) / np.sqrt(n)
    h = g[n * n: n * n + n].astype(np.float32).astype(np.float64) / np.sqrt(n) if field else np.zeros(n)
    return ext_from_upper(np.triu(J, 1), h)


def int_dense(n: int, seed: int, lo: int = -8, hi: int = 8, density: float = 0.5,
              field: bool = True) -> np.ndarray:
    w = splitmix64_np(seed, 2 * n * n + n).astype(np.uint64)
    span = np.uint64(hi - lo + 1)
    vals = ((w[: n * n] >> np.uint64(32)) * span >> np.uint64(32)).astype(np.int64) + lo
    keep = (w[n * n: 2 * n * n] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 < density
    J = np.where(keep, vals, 0).reshape(n, n).astype(float)
    h = ((w[2 * n * n:] >> np.uint64(32)) * span >> np.uint64(32)).astype(np.int64) + lo
    return ext_from_upper(np.triu(J, 1), h.astype(float) if field else np.zeros(n))


def ea3d_csr(L: int, seed: int):
    """3D Edwards-Anderson +-J, periodic L^3, as the full symmetric (n+1)-row
    CSR (sorted columns, no field; corner 1)."""
    n = L ** 3
    w = splitmix64_np(seed, 3 * n)
    sign = np.where((w >> np.uint64(63)) == 1, 1.0, -1.0)
    rows = [dict() for _ in range(n + 1)]
    for i in range(n):
        x, y, z = i % L, (i // L) % L, i // (L * L)
        nbrs = [((x + 1) % L) + L * y + L * L * z, x + L * ((y + 1) % L) + L * L * z,
                x + L * y + L * L * ((z + 1) % L)]
        for d, j in enumerate(nbrs):
            v = sign[3 * i + d]
            rows[i][j] = rows[i].get(j, 0.0) + v
            rows[j][i] = rows[j].get(i, 0.0) + v
    rows[n][n] = 1.0
    rowptr, col, val = [0], [], []
    for i in range(n + 1):
        for j in sorted(rows[i]):
            if rows[i][j] != 0:
                col.append(j)
                val.append(rows[i][j])
        rowptr.append(len(col))
    return n, np.array(rowptr, np.int64), np.array(col, np.int64), np.array(val, float)


def random_pm1(R: int, n: int, seed: int) -> np.ndarray:
    w = splitmix64_np(seed, R * n)
    return np.where((w >> np.uint64(63)) == 1, 1, -1).astype(np.int8).reshape(R, n)


def geometric_betas(bmin: float, bmax: float, L: int) -> np.ndarray:
    """beta_k by repeated multiplication (include/openjij/utility/schedule_list.hpp:165-181)."""
    r = (bmax / bmin) ** (1.0 / (L - 1))
    out = np.empty(L)
    b = bmin
    for k in range(L):
        out[k] = b
        b *= r
    return out


def ks_2samp_pvalue(a, b) -> float:
    from scipy import stats

    return float(stats.ks_2samp(a, b).pvalue)


def qubo_batch_sizes(K: int, seed: int) -> np.ndarray:
    """N_k = round(32 * 32**u), u uniform on [0, 1) from splitmix64: log-uniform
    sizes in [32, 1024]."""
    w = splitmix64_np(seed, K)
    u = (w >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.rint(32.0 * 32.0 ** u).astype(np.int64)


def qubo_upper(n: int, seed: int, lo: int = -8, hi: int = 8, density: float = 0.5) -> np.ndarray:
    """Integer QUBO, entries Q_ij for i <= j present with probability `density`,
    uniform in [lo, hi]; E(x) = sum_{i<=j} Q_ij x_i x_j."""
    w = splitmix64_np(seed, 2 * n * n)
    span = np.uint64(hi - lo + 1)
    vals = ((w[: n * n] >> np.uint64(32)) * span >> np.uint64(32)).astype(np.int64) + lo
    keep = (w[n * n:] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 < density
    return np.triu(np.where(keep, vals, 0).reshape(n, n))


def qubo_to_scaled_ising(Q: np.ndarray):
    """4 x the Ising form of E(x) = sum_{i<=j} Q_ij x_i x_j under x = (1+s)/2:
    J'_ij = Q_ij (i<j), h'_i = 2 Q_ii + sum_{j != i} Q_{min(i,j),max(i,j)},
    so 4 E(x) = E'(s) + C with C = sum_{i<j} Q_ij + 2 sum_i Q_ii.  Integer Q
    stays integer.  Returns (J_ext of E', C)."""
    n = Q.shape[0]
    U = np.triu(Q, 1).astype(float)
    d = np.diag(Q).astype(float)
    h = 2.0 * d + U.sum(axis=1) + U.sum(axis=0)
    return ext_from_upper(U, h), float(U.sum() + 2.0 * d.sum())
