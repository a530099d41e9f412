"""Python bindings (ctypes) for libuvt, the MI355X TRMM engine.

The engine itself is C + HIP (``uvatrmm_amd/csrc``, ABI in ``include/uvt.h``).
This module only loads the in-tree ``lib/libuvt.so`` and wraps its entry
points for the test suite and the benchmark; every computation runs in the
library.  If the library is missing, :func:`lib` raises -- there is no
Python or CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libuvt.so")
CBLAS_PATH = os.path.join(HERE, "lib", "libuvt_cblas.so")
CLI_PATH = os.path.join(HERE, "bin", "uvatrmm.x")

OK, ERR_ARG, ERR_SIZE, ERR_MAPPING, ERR_HIP, ERR_NOMEM, ERR_BUDGET, ERR_UNSUPPORTED, ERR_STATE, \
    ERR_AGAIN = range(10)
MAP_REG, MAP_BALANCED = 0, 1
SHAPE_BOXES, SHAPE_TRAPEZOID = 0, 1
TILE_LOWER, TILE_ACCUM = 1, 2
MAPS = {"reg": MAP_REG, "balanced": MAP_BALANCED}
SHAPES = {"boxes": SHAPE_BOXES, "trapezoid": SHAPE_TRAPEZOID}

_c = ctypes
i32, i64, u32, u64, f64 = _c.c_int32, _c.c_int64, _c.c_uint32, _c.c_uint64, _c.c_double
vp, sz = _c.c_void_p, _c.c_size_t


class Range(_c.Structure):
    _fields_ = [("begin", u64), ("size", u64)]


class Tile(_c.Structure):
    _fields_ = [(k, i32) for k in ("r0", "rows", "c0", "k_lo", "k_hi", "flags", "band", "cfg")]


class Problem(_c.Structure):
    _fields_ = [("A", vp), ("lda", i64), ("B", vp), ("ldb", i64), ("C", vp), ("ldc", i64),
                ("n", i64), ("m", i64), ("prec", i32), ("route", i32)]


class ServeConfig(_c.Structure):
    _fields_ = [("hbm_budget", sz), ("max_slots", i32), ("slices", i32), ("tune_table", _c.c_char_p),
                ("big_flops", f64), ("max_big", i32), ("big_wait_ms", i32)]


class Request(_c.Structure):
    _fields_ = [("n", i64), ("m", i64), ("prec", i32), ("seeded", i32), ("A", vp), ("B", vp),
                ("C", vp), ("seed", u64), ("spread", i32), ("flags", i32), ("tag", u64)]


class Completion(_c.Structure):
    _fields_ = [("ticket", u64), ("tag", u64), ("status", i32), ("path", i32),
                ("t_submit", f64), ("t_admit", f64), ("t_done", f64), ("err_bound", f64),
                ("err_estimate", f64)]


class ServeStats(_c.Structure):
    _fields_ = [("submitted", i64), ("completed", i64), ("rejected", i64), ("p50_s", f64),
                ("p99_s", f64), ("mean_s", f64), ("max_s", f64), ("wall_s", f64), ("flops", f64),
                ("peak_bytes", f64), ("peak_inflight", i32), ("reserved", i32)]


PREC_F64, PREC_SPLIT = 0, 1
ROUTE_AUTO, ROUTE_INT8, ROUTE_FP64 = 0, 1, 2
REQ_AB_DEVICE, REQ_C_DEVICE = 1, 2
TILE_CFGS = {0: "128x128", 1: "64x64", 2: "128x32"}


class UvtError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().uvt_status_string(status).decode() if _LIB else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


_LIB = None

# name -> (restype, argtypes)
_SIGS = {
    "uvt_status_string": (_c.c_char_p, [i32]),
    "uvt_version": (_c.c_char_p, []),
    "uvt_map_band": (i32, [u64, u32, u32, i32, _c.POINTER(Range), _c.POINTER(_c.c_int)]),
    "uvt_plan_bands": (i32, [u64, u32, i32, _c.POINTER(Range), _c.POINTER(_c.c_int)]),
    "uvt_plan_create": (i32, [i64, i64, u32, i32, i32, _c.POINTER(vp), _c.POINTER(_c.c_int)]),
    "uvt_plan_destroy": (i32, [vp]),
    "uvt_plan_phases": (i32, [vp]),
    "uvt_plan_num_tiles": (i64, [vp, i32]),
    "uvt_plan_get_tiles": (i32, [vp, i32, _c.POINTER(Tile), i64]),
    "uvt_plan_get_bands": (i32, [vp, _c.POINTER(Range), u32]),
    "uvt_plan_tile_shape": (i32, [vp, _c.POINTER(i32), _c.POINTER(i32)]),
    "uvt_plan_flops": (f64, [vp]),
    "uvt_plan_issued_flops": (f64, [vp]),
    "uvt_create": (i32, [_c.c_int, _c.POINTER(vp)]),
    "uvt_destroy": (i32, [vp]),
    "uvt_device_cus": (i32, [vp]),
    "uvt_trmm_f64": (i32, [vp, vp, vp, i64, vp, i64, vp, i64, vp]),
    "uvt_plan_launch_info": (i32, [vp, vp, _c.POINTER(i64), _c.POINTER(i64)]),
    "uvt_trmm_split_workspace": (i32, [i64, i64, i32, _c.POINTER(sz)]),
    "uvt_trmm_split": (i32, [vp, i64, i64, i32, vp, i64, vp, i64, vp, i64, vp, sz, vp]),
    "uvt_trmm_split_error": (i32, [vp, i64, i64, i32, vp, sz, _c.POINTER(f64), _c.POINTER(f64), vp]),
    "uvt_batch_create": (i32, [vp, _c.c_void_p, i32, i32, _c.c_char_p, _c.POINTER(vp)]),
    "uvt_batch_run": (i32, [vp, vp]),
    "uvt_batch_update": (i32, [vp, _c.c_void_p, i32]),
    "uvt_batch_info": (i32, [vp, _c.POINTER(f64), _c.POINTER(f64), _c.POINTER(i64),
                             _c.POINTER(i32)]),
    "uvt_batch_errors": (i32, [vp, _c.POINTER(f64), _c.POINTER(f64), i32]),
    "uvt_batch_destroy": (i32, [vp]),
    "uvt_autotune_f64": (i32, [vp, _c.c_char_p]),
    "uvt_tune_choice": (i32, [_c.c_char_p, i64, i64, _c.POINTER(i32)]),
    "uvt_serve_create": (i32, [vp, _c.POINTER(ServeConfig), _c.POINTER(vp)]),
    "uvt_serve_submit": (i32, [vp, _c.POINTER(Request), _c.POINTER(u64)]),
    "uvt_serve_poll": (i32, [vp, _c.POINTER(Completion), i32, _c.POINTER(i32)]),
    "uvt_serve_drain": (i32, [vp]),
    "uvt_serve_stats_get": (i32, [vp, _c.POINTER(ServeStats)]),
    "uvt_serve_request_bytes": (i32, [_c.POINTER(ServeConfig), i64, i64, i32, _c.POINTER(sz)]),
    "uvt_serve_destroy": (i32, [vp]),
    "uvt_row_abs_sums": (i32, [vp, i64, i64, vp, i64, vp, vp]),
    "uvt_inf_norm": (i32, [vp, i64, i64, vp, i64, _c.POINTER(f64), vp]),
    "uvt_fill_closed_form": (i32, [vp, i64, i64, vp, i64, vp, i64, vp]),
    "uvt_fill_seeded": (i32, [vp, i64, i64, u64, i32, vp, i64, vp, i64, vp]),
    "uvt_malloc": (i32, [vp, sz, _c.POINTER(vp)]),
    "uvt_free": (i32, [vp, vp]),
    "uvt_memcpy": (i32, [vp, vp, vp, sz, i32]),
    "uvt_memcpy_2d": (i32, [vp, vp, sz, vp, sz, sz, sz, i32]),
    "uvt_memset": (i32, [vp, vp, i32, sz]),
    "uvt_stream_create": (i32, [vp, _c.POINTER(vp)]),
    "uvt_stream_destroy": (i32, [vp, vp]),
    "uvt_stream_sync": (i32, [vp, vp]),
    "uvt_event_create": (i32, [vp, _c.POINTER(vp)]),
    "uvt_event_destroy": (i32, [vp, vp]),
    "uvt_event_record": (i32, [vp, vp, vp]),
    "uvt_event_elapsed": (i32, [vp, vp, vp, _c.POINTER(_c.c_float)]),
}


def lib():
    """The loaded libuvt.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                               "or make -C uvatrmm_amd")
        L = _c.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(status, what=""):
    if status != OK:
        raise UvtError(status, what)


# ------------------------------------------------------------------ mapping
def map_band(n, p, i, mapping):
    """Band i of p (begin, size) or raises UvtError (with .ref_exit)."""
    r = Range()
    ex = _c.c_int(0)
    st = lib().uvt_map_band(n, p, i, MAPS.get(mapping, mapping), _c.byref(r), _c.byref(ex))
    if st != OK:
        e = UvtError(st, "map_band")
        e.ref_exit = ex.value
        raise e
    return r.begin, r.size


def plan_bands(n, p, mapping):
    arr = (Range * p)()
    ex = _c.c_int(0)
    st = lib().uvt_plan_bands(n, p, MAPS.get(mapping, mapping), arr, _c.byref(ex))
    if st != OK:
        e = UvtError(st, "plan_bands")
        e.ref_exit = ex.value
        raise e
    return [(r.begin, r.size) for r in arr]


class Plan:
    """Tile plan for C = tril(A) * B over p bands (host object)."""

    def __init__(self, n, m, p=1, mapping="reg", shape="trapezoid"):
        self.n, self.m, self.p = n, m, p
        self.mapping, self.shape = mapping, shape
        h = vp()
        ex = _c.c_int(0)
        st = lib().uvt_plan_create(n, m, p, MAPS.get(mapping, mapping), SHAPES.get(shape, shape),
                                   _c.byref(h), _c.byref(ex))
        if st != OK:
            e = UvtError(st, "plan_create")
            e.ref_exit = ex.value
            raise e
        self.ptr = h

    def close(self):
        if self.ptr:
            lib().uvt_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def phases(self):
        return lib().uvt_plan_phases(self.ptr)

    def tiles(self, phase):
        cnt = lib().uvt_plan_num_tiles(self.ptr, phase)
        arr = (Tile * max(cnt, 1))()
        check(lib().uvt_plan_get_tiles(self.ptr, phase, arr, cnt), "get_tiles")
        dt = np.dtype([(k, np.int32) for k in ("r0", "rows", "c0", "k_lo", "k_hi", "flags",
                                               "band", "cfg")])
        return np.frombuffer(bytes(arr), dtype=dt, count=cnt).copy()

    def bands(self):
        arr = (Range * self.p)()
        check(lib().uvt_plan_get_bands(self.ptr, arr, self.p), "get_bands")
        return [(r.begin, r.size) for r in arr]

    def tile_shape(self):
        a, b = i32(), i32()
        check(lib().uvt_plan_tile_shape(self.ptr, _c.byref(a), _c.byref(b)))
        return a.value, b.value

    @property
    def flops(self):
        return lib().uvt_plan_flops(self.ptr)

    @property
    def issued_flops(self):
        return lib().uvt_plan_issued_flops(self.ptr)


# ------------------------------------------------------------------ device
class DeviceBuffer:
    def __init__(self, handle, nbytes):
        self.h = handle
        self.nbytes = nbytes
        p = vp()
        check(lib().uvt_malloc(handle.ptr, nbytes, _c.byref(p)), f"malloc {nbytes}")
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().uvt_memcpy(self.h.ptr, self.ptr, arr.ctypes.data, arr.nbytes, 0), "h2d")

    def download(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().uvt_memcpy(self.h.ptr, out.ctypes.data, self.ptr, out.nbytes, 1), "d2h")
        return out

    def download_rows(self, rows, ncols, ld=None, dtype=np.float64):
        """Selected rows of a row-major matrix with leading dimension ld."""
        ld = ld or ncols
        isz = np.dtype(dtype).itemsize
        out = np.empty((len(rows), ncols), dtype=dtype)
        for k, r in enumerate(rows):
            assert (int(r) * ld + ncols) * isz <= self.nbytes
            check(lib().uvt_memcpy(self.h.ptr, out[k].ctypes.data, self.ptr + int(r) * ld * isz,
                                   ncols * isz, 1), "d2h rows")
        return out

    def upload_2d(self, host, width, height, dpitch, spitch):
        """height rows of width bytes from host (byte pitch spitch) to this
        buffer (byte pitch dpitch)."""
        assert (height - 1) * dpitch + width <= self.nbytes
        assert host.flags.c_contiguous and (height - 1) * spitch + width <= host.nbytes
        check(lib().uvt_memcpy_2d(self.h.ptr, self.ptr, dpitch, host.ctypes.data, spitch, width,
                                  height, 0), "h2d 2d")

    def download_2d(self, host, width, height, dpitch, spitch):
        """height rows of width bytes from this buffer (byte pitch spitch) into
        host (byte pitch dpitch); the rest of host is left as it was."""
        assert (height - 1) * spitch + width <= self.nbytes
        assert host.flags.c_contiguous and (height - 1) * dpitch + width <= host.nbytes
        check(lib().uvt_memcpy_2d(self.h.ptr, host.ctypes.data, dpitch, self.ptr, spitch, width,
                                  height, 1), "d2h 2d")

    def zero(self):
        check(lib().uvt_memset(self.h.ptr, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().uvt_free(self.h.ptr, self.ptr)
            self.ptr = None


class Handle:
    def __init__(self, device=0):
        h = vp()
        check(lib().uvt_create(device, _c.byref(h)), "uvt_create")
        self.ptr = h
        s = vp()
        check(lib().uvt_stream_create(self.ptr, _c.byref(s)), "stream")
        self.stream = s

    @property
    def cus(self):
        return lib().uvt_device_cus(self.ptr)

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def array(self, arr):
        b = DeviceBuffer(self, max(np.asarray(arr).nbytes, 8))
        b.upload(arr)
        return b

    def sync(self):
        check(lib().uvt_stream_sync(self.ptr, self.stream), "sync")

    def trmm_f64(self, plan, A, lda, B, ldb, C, ldc):
        check(lib().uvt_trmm_f64(self.ptr, plan.ptr, A.ptr, lda, B.ptr, ldb, C.ptr, ldc,
                                 self.stream), "trmm_f64")

    def plan_launch_info(self, plan):
        """(workgroups, partial tiles) of the plan's launches on this device."""
        w, p = i64(), i64()
        check(lib().uvt_plan_launch_info(self.ptr, plan.ptr, _c.byref(w), _c.byref(p)), "plan_launch_info")
        return w.value, p.value

    def split_workspace(self, n, m, slices):
        b = sz()
        check(lib().uvt_trmm_split_workspace(n, m, slices, _c.byref(b)), "split_workspace")
        return b.value

    def trmm_split(self, n, m, slices, A, lda, B, ldb, C, ldc, work):
        check(lib().uvt_trmm_split(self.ptr, n, m, slices, A.ptr, lda, B.ptr, ldb, C.ptr, ldc,
                                   work.ptr, work.nbytes, self.stream), "trmm_split")

    def split_error(self, n, m, slices, work):
        """(bound, estimate) of the relative Frobenius error of the last
        trmm_split on `work` (see uvt_trmm_split_error)."""
        b, e = f64(), f64()
        check(lib().uvt_trmm_split_error(self.ptr, n, m, slices, work.ptr, work.nbytes, _c.byref(b),
                                         _c.byref(e), self.stream), "split_error")
        return b.value, e.value

    def inf_norm(self, n, m, C, ldc):
        out = f64()
        check(lib().uvt_inf_norm(self.ptr, n, m, C.ptr, ldc, _c.byref(out), self.stream), "norm")
        return out.value

    def row_abs_sums(self, n, m, C, ldc, out):
        check(lib().uvt_row_abs_sums(self.ptr, n, m, C.ptr, ldc, out.ptr, self.stream))

    def fill_closed_form(self, n, m, A, lda, B, ldb):
        check(lib().uvt_fill_closed_form(self.ptr, n, m, A.ptr if A else None, lda,
                                         B.ptr if B else None, ldb, self.stream), "fill_closed")

    def fill_seeded(self, n, m, seed, spread, A, lda, B, ldb):
        check(lib().uvt_fill_seeded(self.ptr, n, m, seed, spread, A.ptr if A else None, lda,
                                    B.ptr if B else None, ldb, self.stream), "fill_seeded")

    def event(self):
        e = vp()
        check(lib().uvt_event_create(self.ptr, _c.byref(e)))
        return e

    def record(self, ev):
        check(lib().uvt_event_record(self.ptr, ev, self.stream))

    def elapsed_ms(self, e0, e1):
        ms = _c.c_float()
        check(lib().uvt_event_elapsed(self.ptr, e0, e1, _c.byref(ms)))
        return ms.value

    def close(self):
        if self.ptr:
            lib().uvt_stream_destroy(self.ptr, self.stream)
            lib().uvt_destroy(self.ptr)
            self.ptr = None


def tune_choice(n, m, table=None):
    """Tile shape name the tuning table (None: built-in) picks for (n, m)."""
    c = i32()
    check(lib().uvt_tune_choice(table.encode() if table else None, n, m, _c.byref(c)), "tune_choice")
    return TILE_CFGS[c.value]


class Batch:
    """Many TRMMs captured once into a hipGraph (see uvt_batch_create)."""

    def __init__(self, handle, problems, slices=13, table=None):
        self.h = handle
        arr = (Problem * len(problems))(*problems)
        self._arr = arr
        b = vp()
        check(lib().uvt_batch_create(handle.ptr, _c.cast(arr, _c.c_void_p), len(problems), slices,
                                     table.encode() if table else None, _c.byref(b)), "batch_create")
        self.ptr = b
        self.count = len(problems)

    def run(self, stream=None):
        check(lib().uvt_batch_run(self.ptr, stream if stream is not None else self.h.stream),
              "batch_run")

    def update(self, problems):
        arr = (Problem * len(problems))(*problems)
        check(lib().uvt_batch_update(self.ptr, _c.cast(arr, _c.c_void_p), len(problems)),
              "batch_update")
        self._arr = arr

    def info(self):
        fl, i8, tiles, nodes = f64(), f64(), i64(), i32()
        check(lib().uvt_batch_info(self.ptr, _c.byref(fl), _c.byref(i8), _c.byref(tiles),
                                   _c.byref(nodes)))
        return {"flops": fl.value, "int8_flops": i8.value, "fp64_tiles": tiles.value,
                "graph_nodes": nodes.value}

    def errors(self):
        """Per problem (bound, estimate) of the last run; -1 for FP64 ones."""
        b, e = (f64 * self.count)(), (f64 * self.count)()
        check(lib().uvt_batch_errors(self.ptr, b, e, self.count), "batch_errors")
        return list(zip(b, e))

    def close(self):
        if self.ptr:
            lib().uvt_batch_destroy(self.ptr)
            self.ptr = None


def serve_request_bytes(n, m, prec, slices=13):
    """Arena bytes one request of this size and class occupies while in flight."""
    cfg = ServeConfig(1 << 40, 1, slices, None, 0.0, 0, 0)
    b = sz()
    check(lib().uvt_serve_request_bytes(_c.byref(cfg), n, m, prec, _c.byref(b)), "request_bytes")
    return b.value


class Server:
    """Continuous-batching TRMM server (see uvt_serve_create)."""

    def __init__(self, handle, hbm_budget, max_slots=8, slices=13, table=None, big_flops=0.0,
                 max_big=0, big_wait_ms=0):
        self.h = handle
        cfg = ServeConfig(hbm_budget, max_slots, slices, table.encode() if table else None,
                          big_flops, max_big, big_wait_ms)
        p = vp()
        check(lib().uvt_serve_create(handle.ptr, _c.byref(cfg), _c.byref(p)), "serve_create")
        self.ptr = p
        self._host = {}  # ticket -> host arrays the request reads or writes

    def submit_seeded(self, n, m, prec, seed, spread=1, C=None, tag=0):
        """Queues a request whose A and B the device generates from seed.
        C: optional C-contiguous float64 (n, m) array that receives the result."""
        r = Request(n, m, prec, 1, None, None, None, seed, spread, 0, tag)
        keep = []
        if C is not None:
            assert C.dtype == np.float64 and C.flags.c_contiguous and C.shape == (n, m)
            r.C = C.ctypes.data
            keep.append(C)
        return self._submit(r, keep)

    def submit(self, A, B, prec, C=None, tag=0):
        """Queues C = tril(A) @ B for host arrays (copied in on admission)."""
        n, m = B.shape
        assert A.shape == (n, n) and A.dtype == np.float64 and B.dtype == np.float64
        assert A.flags.c_contiguous and B.flags.c_contiguous
        r = Request(n, m, prec, 0, A.ctypes.data, B.ctypes.data, None, 0, 0, 0, tag)
        keep = [A, B]
        if C is not None:
            assert C.dtype == np.float64 and C.flags.c_contiguous and C.shape == (n, m)
            r.C = C.ctypes.data
            keep.append(C)
        return self._submit(r, keep)

    def submit_device(self, n, m, prec, dA, dB, dC, tag=0):
        """Queues C = tril(A) @ B on device-resident operands (DeviceBuffer;
        leading dimensions n, m, m), read and written in place.  dA/dB None:
        The buffers must stay allocated until the request completes."""
        r = Request(n, m, prec, 0, dA.ptr, dB.ptr, dC.ptr, 0, 0,
                    REQ_AB_DEVICE | REQ_C_DEVICE, tag)
        return self._submit(r, [dA, dB, dC])

    def _submit(self, r, keep):
        t = u64()
        check(lib().uvt_serve_submit(self.ptr, _c.byref(r), _c.byref(t)), "serve_submit")
        self._host[t.value] = keep
        return t.value

    def poll(self, cap=4096):
        out = (Completion * cap)()
        k = i32()
        check(lib().uvt_serve_poll(self.ptr, out, cap, _c.byref(k)), "serve_poll")
        res = []
        for c in out[:k.value]:
            self._host.pop(c.ticket, None)
            res.append({"ticket": c.ticket, "tag": c.tag, "status": c.status, "path": c.path,
                        "t_submit": c.t_submit, "t_admit": c.t_admit, "t_done": c.t_done,
                        "err_bound": c.err_bound, "err_estimate": c.err_estimate})
        return res

    def drain(self):
        check(lib().uvt_serve_drain(self.ptr), "serve_drain")

    def stats(self):
        st = ServeStats()
        check(lib().uvt_serve_stats_get(self.ptr, _c.byref(st)), "serve_stats")
        return {k: getattr(st, k) for k, _ in ServeStats._fields_ if k != "reserved"}

    def close(self):
        if self.ptr:
            lib().uvt_serve_destroy(self.ptr)
            self.ptr = None
            self._host.clear()


def autotune(handle, path):
    check(lib().uvt_autotune_f64(handle.ptr, path.encode()), "autotune")


def gpu_available():
    """True when libuvt can open device 0 (never raises)."""
    try:
        h = vp()
        st = lib().uvt_create(0, _c.byref(h))
        if st == OK:
            lib().uvt_destroy(h)
            return True
    except Exception:
        pass
    return False
