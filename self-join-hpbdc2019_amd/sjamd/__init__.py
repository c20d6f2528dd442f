"""Python view of libsjamd, the MI355X grid-indexed epsilon self-join.

Thin ctypes layer over the C ABI in include/sjamd.h.  Device buffers come
from torch (ROCm) tensors; the library does all the work in HIP kernels.
Loading fails loudly when the native library is missing: there is no
CPU fallback in this package.

Reference interface map (benoitgallet/self-join-hpbdc2019):
  Context            warmUpGPU                 GPU.h:54
  Context.grid_build generateNDGridDimensions  main.cu:446-503
                     populateNDGridIndexAndLookupArray  main.cu:240-413
  Grid.join          distanceTableNDGridBatches GPU.h:37-40
  Result.neighbor_table constructNeighborTableKeyValueWithPtrs GPU.cu:2595
  Grid.workload      sortByWorkLoad[LidUnicomp] kernel.cu:33-250
  Grid.batch_estimate callGPUBatchEst[WorkQueue] GPU.cu:854-1284
  import_csv         importNDDataset           import_dataset.cpp:34
  sort_in_nd_bins    sortInNDBins              import_dataset.cpp:76
  Context.many       (no counterpart: many joins in one graph launch)
"""
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsjamd.so")
MAXK = 12

FULL, UNICOMP, BUNICOMP, LID = 0, 1, 2, 3
FP64, FP32_RECHECK = 0, 1
STRIDED, WORKQUEUE, THREADS_PER_POINT = 0, 1, 2

FORMS = {"full": FULL, "unicomp": UNICOMP, "bunicomp": BUNICOMP, "lid": LID}

# reference search modes (main.cu:88-150) -> (form, schedule)
SEARCH_MODES = {
    3: (FULL, STRIDED),
    4: (UNICOMP, STRIDED),
    5: (BUNICOMP, STRIDED),
    6: (LID, STRIDED),
    7: (FULL, WORKQUEUE),
    8: (FULL, THREADS_PER_POINT),
    9: (FULL, WORKQUEUE),
    10: (LID, WORKQUEUE),
}


class SJError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"sjamd status {status}: {msg}")
        self.status = status


class JoinParams(C.Structure):
    _fields_ = [
        ("form", C.c_int32), ("precision", C.c_int32), ("schedule", C.c_int32),
        ("sort_rows", C.c_int32), ("capacity", C.c_uint64), ("block_size", C.c_int32),
        ("queue_chunk", C.c_int32), ("cells_per_tile", C.c_int32), ("hit_record", C.c_int32),
    ]


class GridInfo(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("d", C.c_int32), ("k", C.c_int32), ("eps", C.c_double),
        ("minArr", C.c_double * MAXK), ("maxArr", C.c_double * MAXK),
        ("nCells", C.c_uint32 * MAXK), ("totalCells", C.c_uint64), ("nNonEmpty", C.c_uint64),
        ("dense_lookup", C.c_int32), ("reserved", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("total_pairs", C.c_uint64), ("num_batches", C.c_uint64), ("candidates", C.c_uint64),
        ("fp32_band_pairs", C.c_uint64), ("fp32_misclassified", C.c_uint64),
        ("overflow_retries", C.c_uint64),
        ("ms_grid", C.c_double), ("ms_workload", C.c_double), ("ms_count", C.c_double),
        ("ms_scan", C.c_double), ("ms_fill", C.c_double), ("ms_sort_rows", C.c_double),
        ("ms_total", C.c_double), ("fill_from_record", C.c_uint64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Instance(C.Structure):
    _fields_ = [("points", C.c_void_p), ("n", C.c_uint64), ("d", C.c_int32), ("k", C.c_int32),
                ("eps", C.c_double)]


class ManyTuning(C.Structure):
    _fields_ = [("form", C.c_int32), ("schedule", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("queue_chunk", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class NeighborLookup(C.Structure):
    _fields_ = [("pointID", C.c_int32), ("indexmin", C.c_int32), ("indexmax", C.c_int32),
                ("batch", C.c_int32)]


EXPORTS = [
    "sj_last_error", "sj_context_create", "sj_context_destroy", "sj_context_set_stream",
    "sj_grid_build", "sj_grid_get_info", "sj_grid_copy_arrays", "sj_grid_nd_mask", "sj_grid_free",
    "sj_join", "sj_self_join", "sj_result_get_stats", "sj_result_total_pairs",
    "sj_result_num_points", "sj_result_device_table", "sj_result_device_csr", "sj_result_copy_csr", "sj_result_copy_pairs",
    "sj_result_neighbor_table", "sj_result_batch_data", "sj_result_free", "sj_workload",
    "sj_batch_estimate", "sj_batch_estimate_workqueue", "sj_import_csv", "sj_sort_in_nd_bins", "sj_free_host", "sj_eps_threshold",
    "sj_many_create", "sj_many_run", "sj_many_autotune", "sj_many_set_tuning", "sj_many_get_tuning",
    "sj_many_instance", "sj_many_stats", "sj_many_times", "sj_many_count", "sj_many_free",
    "sj_fp32_band", "sj_mfma_band",
    "sj_self_join_static", "sj_static_need", "sj_static_scratch_bytes",
]

_lib = None


def lib():
    """Load libsjamd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libsjamd.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int32
    P = C.POINTER
    L.sj_last_error.restype = C.c_char_p
    L.sj_eps_threshold.restype = C.c_double
    L.sj_eps_threshold.argtypes = [C.c_double]
    fp = C.POINTER(C.c_float)
    L.sj_fp32_band.restype = i32
    L.sj_fp32_band.argtypes = [C.c_double, C.c_double, i32, fp, fp, fp]
    L.sj_mfma_band.restype = i32
    L.sj_mfma_band.argtypes = [C.c_double, C.c_double, i32, fp, fp, fp, fp]
    L.sj_context_create.argtypes = [i32, vp, P(vp)]
    L.sj_context_destroy.argtypes = [vp]
    L.sj_context_destroy.restype = None
    L.sj_context_set_stream.argtypes = [vp, vp]
    L.sj_grid_build.argtypes = [vp, vp, u64, i32, i32, C.c_double, P(vp)]
    L.sj_grid_get_info.argtypes = [vp, P(GridInfo)]
    L.sj_grid_copy_arrays.argtypes = [vp, vp, vp, vp, vp, vp]
    L.sj_grid_nd_mask.argtypes = [vp, vp, vp, P(u64)]
    L.sj_grid_free.argtypes = [vp]
    L.sj_grid_free.restype = None
    L.sj_join.argtypes = [vp, P(JoinParams), P(vp)]
    L.sj_self_join.argtypes = [vp, vp, u64, i32, i32, C.c_double, P(JoinParams), P(vp)]
    L.sj_result_get_stats.argtypes = [vp, P(Stats)]
    L.sj_result_total_pairs.argtypes = [vp]
    L.sj_result_total_pairs.restype = u64
    L.sj_result_num_points.argtypes = [vp]
    L.sj_result_num_points.restype = u64
    L.sj_result_device_csr.argtypes = [vp, P(vp), P(vp)]
    L.sj_result_device_table.argtypes = [vp, P(vp), P(vp), P(vp)]
    L.sj_result_copy_csr.argtypes = [vp, vp, vp]
    L.sj_result_copy_pairs.argtypes = [vp, vp, vp]
    L.sj_result_neighbor_table.argtypes = [vp, vp]
    L.sj_result_batch_data.argtypes = [vp, u64, P(vp), P(u64)]
    L.sj_result_free.argtypes = [vp]
    L.sj_result_free.restype = None
    L.sj_workload.argtypes = [vp, i32, vp, vp]
    L.sj_batch_estimate.argtypes = [vp, u64, P(u64), P(u64), P(u64), P(u64), P(u64)]
    L.sj_batch_estimate_workqueue.argtypes = [vp, i32, u64, P(u64), P(u64), P(u64), P(u64), P(u64)]
    L.sj_import_csv.argtypes = [C.c_char_p, i32, P(P(C.c_double)), P(u64)]
    L.sj_sort_in_nd_bins.argtypes = [P(C.c_double), u64, i32, i32]
    L.sj_free_host.argtypes = [vp]
    L.sj_free_host.restype = None
    L.sj_many_create.argtypes = [vp, P(Instance), i32, P(JoinParams), P(vp)]
    L.sj_many_run.argtypes = [vp]
    L.sj_many_autotune.argtypes = [vp, i32, i32, P(ManyTuning)]
    L.sj_many_set_tuning.argtypes = [vp, P(ManyTuning)]
    L.sj_many_get_tuning.argtypes = [vp, P(ManyTuning)]
    L.sj_many_instance.argtypes = [vp, i32, P(vp), P(vp), P(vp), P(u64)]
    L.sj_many_stats.argtypes = [vp, i32, P(Stats)]
    L.sj_many_times.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double)]
    L.sj_many_count.argtypes = [vp]
    L.sj_many_count.restype = i32
    L.sj_many_free.argtypes = [vp]
    L.sj_many_free.restype = None
    L.sj_self_join_static.argtypes = [vp, vp, u64, i32, i32, C.c_double, P(JoinParams), vp, vp, u64, vp, P(u64)]
    L.sj_static_need.argtypes = [u64, i32, i32, P(JoinParams)]
    L.sj_static_need.restype = u64
    L.sj_static_scratch_bytes.argtypes = []
    L.sj_static_scratch_bytes.restype = u64
    _lib = L
    return L


def _check(status):
    if status != 0:
        msg = lib().sj_last_error()
        raise SJError(status, msg.decode() if msg else "")


def _np_ptr(a):
    return C.c_void_p(a.ctypes.data)


def eps_threshold(eps):
    return lib().sj_eps_threshold(float(eps))


def fp32_band(T, R, D):
    """(usable, lo, hi, t32) of the VALU FP32 precision form."""
    lo, hi, t32 = C.c_float(), C.c_float(), C.c_float()
    u = lib().sj_fp32_band(float(T), float(R), int(D), C.byref(lo), C.byref(hi), C.byref(t32))
    return bool(u), lo.value, hi.value, t32.value


def mfma_band(T, R, d):
    """(DP, alpha, beta, tlo, thi) of the matrix-core FP32 form."""
    v = [C.c_float() for _ in range(4)]
    dp = lib().sj_mfma_band(float(T), float(R), int(d), *[C.byref(x) for x in v])
    return (dp,) + tuple(x.value for x in v)


_hip_rt = None


def _hip():
    global _hip_rt
    if _hip_rt is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        _hip_rt = h
    return _hip_rt


def download(ptr, count, dtype):
    """Copy `count` elements of a device array to a new numpy array."""
    out = np.empty(count, dtype=dtype)
    if count:
        rc = _hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), out.nbytes, 2)
        if rc != 0:
            raise SJError(3, f"hipMemcpy device->host failed ({rc})")
    return out


def device_copy(ptr, count, torch_dtype, device=0):
    """Copy `count` elements of a device array into a new torch tensor on
    the same device (device to device)."""
    import torch

    out = torch.empty(count, dtype=torch_dtype, device=f"cuda:{device}")
    if count:
        torch.cuda.synchronize()
        rc = _hip().hipMemcpy(C.c_void_p(out.data_ptr()), C.c_void_p(ptr),
                              count * out.element_size(), 3)
        if rc != 0:
            raise SJError(3, f"hipMemcpy device->device failed ({rc})")
    return out


def make_params(form=LID, precision=FP64, schedule=STRIDED, sort_rows=False, capacity=0,
                queue_chunk=0, per_point=False, hit_record=True):
    """per_point=True disables the LDS tile walk (every lane walks its
    neighbour cells from memory); hit_record=False makes the fill evaluate
    every distance again instead of replaying the count pass's hit masks.
    Results are identical either way."""
    if isinstance(form, str):
        form = FORMS[form]
    return JoinParams(form=form, precision=precision, schedule=schedule,
                      sort_rows=1 if sort_rows else 0, capacity=capacity, block_size=0,
                      queue_chunk=queue_chunk, cells_per_tile=-1 if per_point else 0,
                      hit_record=0 if hit_record else -1)


class Context:
    """Device, stream and caching device pool (replaces warmUpGPU)."""

    def __init__(self, device=0, stream=None):
        import torch

        self.torch = torch
        self.device = device
        self.stream = stream  # a torch.cuda.Stream or None
        h = C.c_void_p()
        sp = C.c_void_p(stream.cuda_stream) if stream is not None else C.c_void_p()
        _check(lib().sj_context_create(device, sp, C.byref(h)))
        self.h = h
        # grids, results and batches made on this context: freed before it
        self._children = weakref.WeakSet()

    def _adopt(self, child):
        self._children.add(child)
        return child

    def close(self):
        if self.h:
            for child in list(self._children):
                child.free()
            lib().sj_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def to_device(self, points):
        """float64 (n, d) array -> contiguous device tensor.  A conversion or
        copy runs on torch's current stream; when that is not the context's
        stream it is waited for here, so the library never reads the
        coordinates before they are written."""
        torch = self.torch
        if isinstance(points, torch.Tensor):
            t = points
        else:
            t = torch.from_numpy(np.ascontiguousarray(points, dtype=np.float64))
        out = t.to(device=f"cuda:{self.device}", dtype=torch.float64).contiguous()
        if out is not t or out.data_ptr() != t.data_ptr():
            cur = torch.cuda.current_stream(self.device)
            if self.stream is None or cur.cuda_stream != self.stream.cuda_stream:
                cur.synchronize()
        return out

    def grid_build(self, points, k, eps):
        t = self.to_device(points)
        n, d = t.shape
        h = C.c_void_p()
        _check(lib().sj_grid_build(self.h, C.c_void_p(t.data_ptr()), n, d, k, float(eps),
                                   C.byref(h)))
        return self._adopt(Grid(self, h, t))

    def self_join(self, points, k, eps, **kw):
        t = self.to_device(points)
        n, d = t.shape
        p = make_params(**kw)
        h = C.c_void_p()
        _check(lib().sj_self_join(self.h, C.c_void_p(t.data_ptr()), n, d, k, float(eps),
                                  C.byref(p), C.byref(h)))
        return self._adopt(Result(self, h))

    def many(self, instances, **kw):
        """Plan many joins; instances: iterable of (points, k, eps)."""
        return self._adopt(Many(self, instances, **kw))


class Grid:
    def __init__(self, ctx, h, tensor):
        self.ctx, self.h, self._t = ctx, h, tensor
        gi = GridInfo()
        _check(lib().sj_grid_get_info(h, C.byref(gi)))
        self.info = gi
        self.n, self.d, self.k = gi.n, gi.d, gi.k
        self.nNonEmpty = gi.nNonEmpty

    def free(self):
        if self.h:
            lib().sj_grid_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def dims(self):
        k = self.k
        gi = self.info
        return (np.array(gi.minArr[:k]), np.array(gi.maxArr[:k]),
                np.array(gi.nCells[:k], dtype=np.uint32), gi.totalCells)

    def arrays(self):
        nne, n, k = self.nNonEmpty, self.n, self.k
        lin = np.zeros(nne, dtype=np.uint64)
        imin = np.zeros(nne, dtype=np.int32)
        imax = np.zeros(nne, dtype=np.int32)
        look = np.zeros(n, dtype=np.uint32)
        coord = np.zeros(nne * k, dtype=np.uint32)
        _check(lib().sj_grid_copy_arrays(self.h, _np_ptr(lin), _np_ptr(imin), _np_ptr(imax),
                                         _np_ptr(look), _np_ptr(coord)))
        return dict(cellLinear=lin, indexMin=imin, indexMax=imax, indexLookup=look,
                    cellCoord=coord.reshape(nne, k))

    def nd_mask(self):
        ne = C.c_uint64()
        _check(lib().sj_grid_nd_mask(self.h, None, None, C.byref(ne)))
        mask = np.zeros(ne.value, dtype=np.uint32)
        offs = np.zeros(2 * self.k, dtype=np.uint32)
        _check(lib().sj_grid_nd_mask(self.h, _np_ptr(mask), _np_ptr(offs), C.byref(ne)))
        return mask, offs

    def join(self, **kw):
        p = make_params(**kw)
        h = C.c_void_p()
        _check(lib().sj_join(self.h, C.byref(p), C.byref(h)))
        return self.ctx._adopt(Result(self.ctx, h))

    def workload(self, lid=False):
        w = np.zeros(self.nNonEmpty, dtype=np.int64)
        o = np.zeros(self.nNonEmpty, dtype=np.uint32)
        _check(lib().sj_workload(self.h, 1 if lid else 0, _np_ptr(w), _np_ptr(o)))
        return w, o

    def batch_estimate(self, stride=100, workqueue=None):
        """callGPUBatchEst (workqueue=None: every stride-th point) or
        callGPUBatchEstWorkQueue (workqueue='full' or 'lid': the first
        n/stride points in that workload order)."""
        vals = [C.c_uint64() for _ in range(5)]
        if workqueue is None:
            _check(lib().sj_batch_estimate(self.h, stride, *[C.byref(v) for v in vals]))
        else:
            lid = {"full": 0, "lid": 1}[workqueue]
            _check(lib().sj_batch_estimate_workqueue(self.h, lid, stride, *[C.byref(v) for v in vals]))
        cnt, est, esta, nb, buf = (v.value for v in vals)
        return dict(sample_count=cnt, estimated=est, estimated_with_alpha=esta, num_batches=nb,
                    buffer_size=buf)


class Result:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        self.total = lib().sj_result_total_pairs(h)
        self.n = lib().sj_result_num_points(h)

    def free(self):
        if self.h:
            lib().sj_result_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def stats(self):
        s = Stats()
        _check(lib().sj_result_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def device_csr(self):
        o, i = C.c_void_p(), C.c_void_p()
        _check(lib().sj_result_device_csr(self.h, C.byref(o), C.byref(i)))
        return o.value, i.value

    def device_table(self):
        o, i, r = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().sj_result_device_table(self.h, C.byref(o), C.byref(i), C.byref(r)))
        return o.value, i.value, r.value

    def csr(self):
        off = np.zeros(self.n + 1, dtype=np.uint64)
        ids = np.zeros(max(self.total, 1), dtype=np.int32)
        _check(lib().sj_result_copy_csr(self.h, _np_ptr(off), _np_ptr(ids)))
        return off, ids[:self.total]

    def pairs(self):
        k = np.zeros(max(self.total, 1), dtype=np.int32)
        v = np.zeros(max(self.total, 1), dtype=np.int32)
        _check(lib().sj_result_copy_pairs(self.h, _np_ptr(k), _np_ptr(v)))
        return k[:self.total], v[:self.total]

    def neighbor_table(self):
        tab = np.empty((self.n, 4), dtype=np.int32)  # every entry is written
        _check(lib().sj_result_neighbor_table(self.h, _np_ptr(tab)))
        return tab

    def batch_data(self, b):
        p, cnt = C.c_void_p(), C.c_uint64()
        _check(lib().sj_result_batch_data(self.h, b, C.byref(p), C.byref(cnt)))
        if cnt.value == 0:
            return np.zeros(0, dtype=np.int32)
        arr = (C.c_int32 * cnt.value).from_address(p.value)
        return np.frombuffer(arr, dtype=np.int32).copy()


def import_csv(path, d):
    ptr = C.POINTER(C.c_double)()
    n = C.c_uint64()
    _check(lib().sj_import_csv(str(path).encode(), d, C.byref(ptr), C.byref(n)))
    cnt = n.value * d
    out = np.ctypeslib.as_array(ptr, (max(cnt, 1),))[:cnt].copy().reshape(n.value, d)
    lib().sj_free_host(C.cast(ptr, C.c_void_p))
    return out


def sort_in_nd_bins(points, stable=False):
    p = np.ascontiguousarray(points, dtype=np.float64).copy()
    _check(lib().sj_sort_in_nd_bins(p.ctypes.data_as(C.POINTER(C.c_double)), p.shape[0],
                                    p.shape[1], 1 if stable else 0))
    return p


class Many:
    """Many independent self-joins in one persistent, graph-recorded launch
    per pass (sj_many_* in include/sjamd.h)."""

    def __init__(self, ctx, instances, **kw):
        self.ctx = ctx
        self._t = [ctx.to_device(p) for p, _, _ in instances]
        arr = (Instance * len(self._t))()
        self.shapes = []
        for i, (t, (_, k, eps)) in enumerate(zip(self._t, instances)):
            n, d = t.shape
            arr[i] = Instance(C.c_void_p(t.data_ptr()), n, d, k, float(eps))
            self.shapes.append((n, d, k, float(eps)))
        p = make_params(**kw)
        h = C.c_void_p()
        _check(lib().sj_many_create(ctx.h, arr, len(self._t), C.byref(p), C.byref(h)))
        self.h = h

    def free(self):
        if self.h:
            lib().sj_many_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return len(self.shapes)

    def run(self):
        _check(lib().sj_many_run(self.h))

    def autotune(self, tune_form=False, reps=3):
        t = ManyTuning()
        _check(lib().sj_many_autotune(self.h, 1 if tune_form else 0, reps, C.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = ManyTuning()
        _check(lib().sj_many_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, form, schedule, blocks_per_cu=8, queue_chunk=1):
        if isinstance(form, str):
            form = FORMS[form]
        t = ManyTuning(form, schedule, blocks_per_cu, queue_chunk)
        _check(lib().sj_many_set_tuning(self.h, C.byref(t)))

    def times(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().sj_many_times(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(ms_plan=a.value, ms_grid=b.value, ms_run=c.value)

    def stats(self, i=-1):
        s = Stats()
        _check(lib().sj_many_stats(self.h, i, C.byref(s)))
        return s.as_dict()

    def device_table(self, i):
        o, d, r, tot = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_uint64()
        _check(lib().sj_many_instance(self.h, i, C.byref(o), C.byref(d), C.byref(r), C.byref(tot)))
        return o.value, d.value, r.value, tot.value

    def csr(self, i):
        """Host CSR of instance i in original point-ID order."""
        n = self.shapes[i][0]
        o, d, r, tot = self.device_table(i)
        off = download(o, n + 1, np.uint64).astype(np.int64)
        ids = download(d, tot, np.int32) if tot else np.zeros(0, dtype=np.int32)
        rows = download(r, n, np.uint32).astype(np.int64)
        cnt = np.diff(off)
        out_cnt = np.zeros(n, dtype=np.int64)
        out_cnt[rows] = cnt
        out_off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(out_cnt, out=out_off[1:])
        out = np.empty(tot, dtype=np.int32)
        # row r (cell order) of point rows[r] moves to out_off[rows[r]]
        src = np.repeat(off[:-1], cnt)
        dst = np.repeat(out_off[rows], cnt)
        within = np.arange(tot, dtype=np.int64) - np.repeat(off[:-1], cnt)
        out[dst + within] = ids[src + within]
        return out_off.astype(np.uint64), out

