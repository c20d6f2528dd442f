"""ctypes view of the CPU oracle (libsjoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
benchmark's cpu_baseline leg.  The product package never imports it.

The C restatement it wraps (sj_oracle.c) cites the reference file:line of
every function.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsjoracle.so")
_lib = None

FULL, UNICOMP, BUNICOMP, LID = 0, 1, 2, 3
MAXK = 12


class Grid(C.Structure):
    _fields_ = [
        ("d", C.c_int), ("k", C.c_int), ("eps", C.c_double), ("n", C.c_uint64),
        ("minArr", C.c_double * MAXK), ("maxArr", C.c_double * MAXK),
        ("nCells", C.c_uint32 * MAXK), ("totalCells", C.c_uint64),
        ("nNonEmpty", C.c_uint64),
        ("cellLinear", C.POINTER(C.c_uint64)),
        ("indexMin", C.POINTER(C.c_int32)), ("indexMax", C.POINTER(C.c_int32)),
        ("indexLookup", C.POINTER(C.c_uint32)),
        ("cellCoord", C.POINTER(C.c_uint32)),
        ("ndMask", C.POINTER(C.c_uint32)),
        ("ndMaskOffsets", C.c_uint32 * (2 * MAXK)),
        ("nNDMaskElems", C.c_uint64),
    ]


class NTL(C.Structure):
    _fields_ = [("pointID", C.c_int32), ("indexmin", C.c_int32),
                ("indexmax", C.c_int32), ("batch", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        u64p = C.POINTER(C.c_uint64)
        L.sjo_eps_threshold.restype = C.c_double
        L.sjo_eps_threshold.argtypes = [C.c_double]
        L.sjo_dist2.restype = C.c_double
        L.sjo_dist2.argtypes = [dp, dp, C.c_int]
        L.sjo_dist2_batch.argtypes = [dp, dp, C.c_uint64, C.c_int, dp]
        L.sjo_linear_id.restype = C.c_uint64
        L.sjo_grid_dims.argtypes = [dp, C.c_uint64, C.c_int, C.c_int, C.c_double,
                                    dp, dp, C.POINTER(C.c_uint32), u64p]
        L.sjo_grid_build.argtypes = [dp, C.c_uint64, C.c_int, C.c_int, C.c_double,
                                     C.POINTER(Grid)]
        L.sjo_grid_free.argtypes = [C.POINTER(Grid)]
        L.sjo_claims.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_int), C.POINTER(C.c_uint32)]
        L.sjo_join.argtypes = [C.POINTER(Grid), dp, C.c_int, C.c_int, C.c_int,
                               u64p, C.POINTER(C.POINTER(C.c_int32))]
        L.sjo_brute.argtypes = [dp, C.c_uint64, C.c_int, C.c_double, C.c_int,
                                u64p, C.POINTER(C.POINTER(C.c_int32))]
        L.sjo_workload.argtypes = [C.POINTER(Grid), C.c_int, C.POINTER(C.c_int64)]
        L.sjo_workload_order.argtypes = [C.POINTER(Grid), C.c_int, C.POINTER(C.c_uint32)]
        L.sjo_sample_count.restype = C.c_uint64
        L.sjo_sample_count.argtypes = [C.POINTER(Grid), dp, C.c_uint64]
        L.sjo_sample_count_workqueue.restype = C.c_uint64
        L.sjo_sample_count_workqueue.argtypes = [C.POINTER(Grid), dp, C.c_uint64, C.c_int]
        L.sjo_batch_policy.argtypes = [C.c_uint64, C.c_uint64, u64p, u64p, u64p, u64p]
        L.sjo_neighbor_table.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.c_uint64, C.c_int32, C.POINTER(NTL)]
        L.sjo_import_csv.argtypes = [C.c_char_p, C.c_int, C.POINTER(dp), u64p]
        L.sjo_sort_in_nd_bins_stable.argtypes = [dp, C.c_uint64, C.c_int]
        L.sjo_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pts(points):
    p = np.ascontiguousarray(points, dtype=np.float64)
    if p.ndim != 2:
        raise ValueError("points must be (n, d)")
    return p


def eps_threshold(eps):
    return lib().sjo_eps_threshold(float(eps))


def dist2(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().sjo_dist2(_dptr(a), _dptr(b), a.size)


def dist2_batch(a, b):
    """Predicate P's squared distance of each row pair of a and b (m, d)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty(a.shape[0])
    lib().sjo_dist2_batch(_dptr(a), _dptr(b), a.shape[0], a.shape[1], _dptr(out))
    return out


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with status {rc}")


def grid_dims(points, k, eps):
    p = _pts(points)
    mn = np.zeros(MAXK); mx = np.zeros(MAXK)
    nc = np.zeros(MAXK, dtype=np.uint32); tot = C.c_uint64()
    _check(lib().sjo_grid_dims(_dptr(p), p.shape[0], p.shape[1], k, float(eps), _dptr(mn),
                               _dptr(mx), nc.ctypes.data_as(C.POINTER(C.c_uint32)),
                               C.byref(tot)), "grid_dims")
    return mn[:k].copy(), mx[:k].copy(), nc[:k].copy(), tot.value


class OracleGrid:
    """Owns an sjo_grid; exposes its arrays as numpy copies."""

    def __init__(self, points, k, eps):
        self.points = _pts(points)
        self.g = Grid()
        _check(lib().sjo_grid_build(_dptr(self.points), self.points.shape[0],
                                    self.points.shape[1], k, float(eps), C.byref(self.g)),
               "grid_build")
        g = self.g
        nne = g.nNonEmpty
        self.k, self.d, self.eps, self.n = g.k, g.d, g.eps, g.n
        self.minArr = np.array(g.minArr[:k])
        self.maxArr = np.array(g.maxArr[:k])
        self.nCells = np.array(g.nCells[:k], dtype=np.uint32)
        self.totalCells = g.totalCells
        self.nNonEmpty = nne
        self.cellLinear = np.ctypeslib.as_array(g.cellLinear, (nne,)).copy()
        self.indexMin = np.ctypeslib.as_array(g.indexMin, (nne,)).copy()
        self.indexMax = np.ctypeslib.as_array(g.indexMax, (nne,)).copy()
        self.indexLookup = np.ctypeslib.as_array(g.indexLookup, (g.n,)).copy()
        self.cellCoord = np.ctypeslib.as_array(g.cellCoord, (nne * k,)).copy().reshape(nne, k)
        self.ndMask = np.ctypeslib.as_array(g.ndMask, (g.nNDMaskElems,)).copy()
        self.ndMaskOffsets = np.array(g.ndMaskOffsets[:2 * k], dtype=np.uint32)

    def __del__(self):
        try:
            lib().sjo_grid_free(C.byref(self.g))
        except Exception:
            pass

    def join(self, form=LID, nthreads=0, sort_rows=True):
        n = self.n
        off = np.zeros(n + 1, dtype=np.uint64)
        ids = C.POINTER(C.c_int32)()
        _check(lib().sjo_join(C.byref(self.g), _dptr(self.points), form, nthreads,
                              1 if sort_rows else 0,
                              off.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(ids)), "join")
        total = int(off[-1])
        out = np.ctypeslib.as_array(ids, (max(total, 1),))[:total].copy()
        lib().sjo_free(ids)
        return off, out

    def workload(self, lid=False):
        w = np.zeros(self.nNonEmpty, dtype=np.int64)
        _check(lib().sjo_workload(C.byref(self.g), 1 if lid else 0,
                                  w.ctypes.data_as(C.POINTER(C.c_int64))), "workload")
        return w

    def workload_order(self, lid=False):
        o = np.zeros(self.nNonEmpty, dtype=np.uint32)
        _check(lib().sjo_workload_order(C.byref(self.g), 1 if lid else 0,
                                        o.ctypes.data_as(C.POINTER(C.c_uint32))), "order")
        return o

    def sample_count(self, stride=100):
        return lib().sjo_sample_count(C.byref(self.g), _dptr(self.points), stride)

    def sample_count_workqueue(self, stride=100, lid=False):
        return lib().sjo_sample_count_workqueue(C.byref(self.g), _dptr(self.points), stride,
                                                1 if lid else 0)


def join(points, k, eps, form=LID, nthreads=0, sort_rows=True):
    return OracleGrid(points, k, eps).join(form, nthreads, sort_rows)


def brute(points, eps, nthreads=0):
    p = _pts(points)
    n = p.shape[0]
    off = np.zeros(n + 1, dtype=np.uint64)
    ids = C.POINTER(C.c_int32)()
    _check(lib().sjo_brute(_dptr(p), n, p.shape[1], float(eps), nthreads,
                           off.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(ids)), "brute")
    total = int(off[-1])
    out = np.ctypeslib.as_array(ids, (max(total, 1),))[:total].copy()
    lib().sjo_free(ids)
    return off, out


def claims(form, k, cell, off, nCells):
    c = np.ascontiguousarray(cell, dtype=np.uint32)
    o = np.ascontiguousarray(off, dtype=np.int32)
    nc = np.ascontiguousarray(nCells, dtype=np.uint32)
    return bool(lib().sjo_claims(form, k, c.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 o.ctypes.data_as(C.POINTER(C.c_int)),
                                 nc.ctypes.data_as(C.POINTER(C.c_uint32))))


def batch_policy(sample_count, offset_rate=100):
    vals = [C.c_uint64() for _ in range(4)]
    lib().sjo_batch_policy(int(sample_count), int(offset_rate), *[C.byref(v) for v in vals])
    est, est_alpha, nb, buf = (v.value for v in vals)
    return {"estimated": est, "estimated_with_alpha": est_alpha,
            "num_batches": nb, "buffer_size": buf}


def neighbor_table(keys, vals, n, batch=0):
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.int32)
    tab = (NTL * n)()
    for i in range(n):
        tab[i].pointID = -1
        tab[i].indexmin = 0
        tab[i].indexmax = -1
        tab[i].batch = -1
    lib().sjo_neighbor_table(keys.ctypes.data_as(C.POINTER(C.c_int32)),
                             vals.ctypes.data_as(C.POINTER(C.c_int32)), keys.size, batch, tab)
    return np.array([(t.pointID, t.indexmin, t.indexmax, t.batch) for t in tab],
                    dtype=np.int64).reshape(n, 4)


def import_csv(path, d):
    ptr = C.POINTER(C.c_double)()
    n = C.c_uint64()
    _check(lib().sjo_import_csv(path.encode(), d, C.byref(ptr), C.byref(n)), "import_csv")
    cnt = n.value * d
    out = np.ctypeslib.as_array(ptr, (max(cnt, 1),))[:cnt].copy().reshape(n.value, d)
    lib().sjo_free(ptr)
    return out


def sort_in_nd_bins_stable(points):
    p = _pts(points).copy()
    _check(lib().sjo_sort_in_nd_bins_stable(_dptr(p), p.shape[0], p.shape[1]), "sort_in_nd_bins")
    return p


def csr_rows_sorted(offsets, ids):
    """Canonical form: each row sorted ascending (rows stay in place)."""
    ids = ids.copy()
    n = offsets.size - 1
    row = np.repeat(np.arange(n, dtype=np.int64), np.diff(offsets.astype(np.int64)))
    order = np.lexsort((ids, row))
    return offsets, ids[order]
