"""Deterministic synthetic point sets for the self-join workloads.

Every coordinate comes from an integer-defined SplitMix64 stream, so the same
seed gives the same bits on any machine with IEEE float64 numpy:
u = (x >> 11) * 2**-53 in [0, 1).

Workloads (sizes as benchmarked):
  uniform2d   100,000 x 2-D, U[0,100]^2, eps 0.5
  exp2d       2,000,000 x 2-D, Exp(lambda=40) clipped to [0,1], mean 64 neighbours
  uniform6d   2,000,000 x 6-D, U[0,1]^6, mean 32 neighbours
  clusters64d 500,000 x 64-D, 100 Gaussian clusters (centres U[0,10]^64,
              sigma 0.5), mean 50 neighbours, 6 indexed dimensions
  multi       64 joins, N log-uniform in [1e4, 5e5], d in {2,3,4,6,8},
              uniform or Exp(40) data, mean neighbours U[8,128]
"""
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, start=0):
    """n outputs of the SplitMix64 generator seeded with `seed`."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
             + (np.arange(start + 1, start + n + 1, dtype=np.uint64) * _GOLDEN))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, start=0):
    return (splitmix64(seed, n, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(n, d, seed, lo=0.0, hi=1.0):
    u = uniform01(seed, n * d).reshape(n, d)
    return lo + (hi - lo) * u


def exponential_clipped(n, d, seed, lam=40.0):
    u = uniform01(seed, n * d).reshape(n, d)
    x = -np.log1p(-u) / lam
    return np.minimum(x, 1.0)


def gaussian(n, d, seed):
    """Box-Muller normals from two SplitMix64 streams."""
    u1 = uniform01(seed, n * d)
    u2 = uniform01(seed ^ 0x5DEECE66D, n * d)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return (r * np.cos(2.0 * math.pi * u2)).reshape(n, d)


def clusters(n, d, seed, n_clusters=100, box=10.0, sigma=0.5):
    centres = uniform(n_clusters, d, seed * 1000 + 1, 0.0, box)
    lab = (uniform01(seed * 1000 + 2, n) * n_clusters).astype(np.int64)
    lab = np.minimum(lab, n_clusters - 1)
    return centres[lab] + sigma * gaussian(n, d, seed * 1000 + 3)


# eps values calibrated by calibrate_eps() (mean neighbour count including
# the point itself over a 20,000-query subsample), stored as exact floats.
CALIBRATED_EPS = {
    "exp2d": "0x1.4e4d7ccc6312cp-13",        # ~1.5941e-4
    "uniform6d": "0x1.ff57925fc924ep-4",    # ~0.12484
    "clusters64d": "0x1.20f5f0245249bp+2",  # ~4.5150
}

WORKLOADS = {
    "uniform2d": dict(n=100_000, d=2, k=2, seed=1, mean=None),
    "exp2d": dict(n=2_000_000, d=2, k=2, seed=2, mean=64),
    "uniform6d": dict(n=2_000_000, d=6, k=6, seed=3, mean=32),
    "clusters64d": dict(n=500_000, d=64, k=6, seed=4, mean=50),
}


def make_points(name, n=None):
    w = WORKLOADS[name]
    n = w["n"] if n is None else n
    if name == "uniform2d":
        return uniform(n, 2, w["seed"], 0.0, 100.0)
    if name == "exp2d":
        return exponential_clipped(n, 2, w["seed"])
    if name == "uniform6d":
        return uniform(n, 6, w["seed"])
    if name == "clusters64d":
        return clusters(n, 64, w["seed"])
    raise KeyError(name)


def workload_eps(name):
    if name == "uniform2d":
        return 0.5
    e = CALIBRATED_EPS[name]
    if e is None:
        raise RuntimeError(f"eps for {name} not calibrated")
    return float.fromhex(e)


def _query_subsample(n, n_query, seed=7):
    return np.unique((uniform01(seed, n_query) * n).astype(np.int64))


def neighbour_profile(points, kmax, n_query=20_000, seed=7):
    """Sorted distances from a fixed query subsample to its kmax nearest
    points (the query itself included).  kd-tree for low d, blocked
    norm-expansion distances for high d (calibration only)."""
    n, d = points.shape
    q = _query_subsample(n, n_query if d <= 8 else min(n_query, 2_000), seed)
    kmax = min(kmax, n)
    if d <= 8:
        from scipy.spatial import cKDTree

        dist, _ = cKDTree(points).query(points[q], k=kmax)
        return np.asarray(dist).reshape(len(q), kmax)
    pn = np.einsum("ij,ij->i", points, points)
    out = np.empty((len(q), kmax))
    for s in range(0, len(q), 100):
        qq = points[q[s:s + 100]]
        d2 = pn[None, :] + np.einsum("ij,ij->i", qq, qq)[:, None] - 2.0 * (qq @ points.T)
        np.maximum(d2, 0.0, out=d2)
        part = np.partition(d2, kmax - 1, axis=1)[:, :kmax]
        out[s:s + 100] = np.sqrt(np.sort(part, axis=1))
    return out


def mean_neighbours(points, eps, n_query=20_000, seed=7, profile=None, kmax=1024):
    """Mean neighbour count (including self) of the fixed query subsample."""
    prof = neighbour_profile(points, kmax, n_query, seed) if profile is None else profile
    return float(np.mean(np.sum(prof <= eps, axis=1)))


def calibrate_eps(points, mean, iters=200, rel=0.01):
    """Deterministic bisection on eps for the given mean neighbour count."""
    prof = neighbour_profile(points, max(1024, int(16 * mean)))
    lo, hi = float(prof[:, 1:].min()) * 0.5, float(prof.max())
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        m = mean_neighbours(points, mid, profile=prof)
        if abs(m - mean) <= rel * mean:
            return mid
        if m < mean:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def multi_join_instances(count=64, seed=5):
    """Instance descriptors of the batched workload (sizes, dims, data kind,
    target mean neighbours); points come from instance_points()."""
    u = uniform01(seed, count * 4).reshape(count, 4)
    dims = np.array([2, 3, 4, 6, 8])
    out = []
    for i in range(count):
        n = int(round(math.exp(math.log(1e4) + u[i, 0] * (math.log(5e5) - math.log(1e4)))))
        d = int(dims[min(int(u[i, 1] * 5), 4)])
        kind = "uniform" if u[i, 2] < 0.5 else "exp"
        mean = 8.0 + u[i, 3] * 120.0
        out.append(dict(n=n, d=d, kind=kind, mean=mean, seed=seed * 100000 + i))
    return out


# eps of the 64 instances of multi_join_instances(), calibrated by
# calibrate_eps() for each instance's mean neighbour count, exact floats.
MULTI_EPS = [
    "0x1.c7ac2079d3230p-3", "0x1.441015ec6cbfep-8", "0x1.cbabd26821592p-3",
    "0x1.4a5c01e31fcd8p-8", "0x1.5ec56f999a254p-4", "0x1.0ddc06dfddc34p-1",
    "0x1.d9927b589bdf0p-7", "0x1.6fb12e2c63961p-2", "0x1.ea7938c6a4d76p-10",
    "0x1.41a61a73223e2p-4", "0x1.da5ebe7ccde80p-3", "0x1.099bd2d4c3714p-2",
    "0x1.e58ac1f9a7681p-4", "0x1.e8b02f6d36852p-7", "0x1.bd9ab3ec121f9p-7",
    "0x1.26161c3722ae6p-6", "0x1.42623ee446ab0p-4", "0x1.7295695a76d6cp-6",
    "0x1.23b34847eb63fp-12", "0x1.4f3d7076d62b8p-8", "0x1.9189f3776c8d4p-9",
    "0x1.0518bd6b5c803p-11", "0x1.71e197df372c0p-6", "0x1.dfa87845775a8p-10",
    "0x1.f939c8fc37dbep-10", "0x1.a33ae88560d14p-8", "0x1.1e9a9086bda64p-8",
    "0x1.01b86517b7474p-7", "0x1.7bbabb2ad7b9fp-8", "0x1.3584a739f802ap-2",
    "0x1.3a38250b09d81p-3", "0x1.b8600d695bbb6p-3", "0x1.9b7d67e0c9272p-2",
    "0x1.ad9998c8c44a8p-5", "0x1.15b98acf43ecdp-9", "0x1.d6f17a8da50e6p-3",
    "0x1.af04fb65d5e70p-12", "0x1.f6dd8b53db663p-5", "0x1.3f8f3fc1065d0p-4",
    "0x1.f9c66e33b2019p-12", "0x1.08100abb0161cp-3", "0x1.1b575db728118p-7",
    "0x1.fd0bcec26322bp-8", "0x1.b6f21e606a7c8p-7", "0x1.5028496e9d450p-2",
    "0x1.a37439b500205p-7", "0x1.411c723dc04afp-10", "0x1.81ff2fbc23542p-7",
    "0x1.1dfce4adb1d86p-6", "0x1.9dc0670ec4225p-5", "0x1.89f8258ad3ed7p-2",
    "0x1.0bf33a30500b5p-2", "0x1.a945c7df6b6c8p-7", "0x1.09934cab221f1p-9",
    "0x1.2fab88d62ece4p-2", "0x1.53d565431e040p-8", "0x1.94005430b77bcp-2",
    "0x1.1e73d67c6c04bp-2", "0x1.6167a222af96ap-8", "0x1.ceb247a594c42p-11",
    "0x1.7434d168536c1p-7", "0x1.6aeb644ff0229p-8", "0x1.a6051f73f88c4p-3",
    "0x1.d870e00ac308ap-6",
]


def multi_eps(i):
    return float.fromhex(MULTI_EPS[i])


def instance_points(inst):
    if inst["kind"] == "uniform":
        return uniform(inst["n"], inst["d"], inst["seed"])
    return exponential_clipped(inst["n"], inst["d"], inst["seed"])
