"""apspie_amd -- MI355X (gfx950) SpMSpV traversal and hash-SpGEMM library.

Python binding over the C ABI in include/apspie.h (libapspie.so, built
in-tree by apspie_amd/build.py).  Host arrays are numpy; device memory is
owned by the library (or passed in as raw device addresses).  There is no
CPU fallback: if the shared library is missing or fails to load, every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libapspie.so")

SSSP_INF = np.float32(1.70141e38)

RMAT_UNDIRECTED = 1
RMAT_WEIGHTED = 2
LOAD_UNDIRECTED = 1
SR_PLUS_TIMES, SR_MIN_PLUS, SR_MAX_TIMES, SR_OR_AND = 0, 1, 2, 3
LOAD_SORT_COLUMNS = 2
LOAD_MM_BANNER = 4

ERRORS = {
    -1: "EINVAL",
    -2: "ENOMEM",
    -3: "EHIP",
    -4: "EIO",
    -5: "ELIMIT",
    -6: "ENOCONV",
    -7: "EBUSY",
}


class ApspieError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Csr(C.Structure):
    _fields_ = [
        ("m", C.c_int32),
        ("n", C.c_int32),
        ("nnz", C.c_int64),
        ("row_ptr", C.c_void_p),
        ("col_ind", C.c_void_p),
        ("val", C.c_void_p),
    ]


class _DCsr(C.Structure):
    """apspie_csr / apspie_csr_owned: device CSR plus the optional CSC."""

    _fields_ = [
        ("m", C.c_int32),
        ("n", C.c_int32),
        ("nnz", C.c_int64),
        ("row_ptr", C.c_void_p),
        ("col_ind", C.c_void_p),
        ("val", C.c_void_p),
        ("csc_col_ptr", C.c_void_p),
        ("csc_row_ind", C.c_void_p),
        ("csc_val", C.c_void_p),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("bfs_alpha", C.c_int32),
        ("bfs_beta", C.c_int32),
        ("sssp_pull_ratio", C.c_int32),
        ("sssp_bin_ratio", C.c_int32),
        ("expand_wg_per_cu", C.c_int32),
        ("round_chunk", C.c_int32),
        ("batch_chunk", C.c_int32),
        ("sssp_batch_lanes", C.c_int32),
        ("sssp_delta", C.c_int32),
        ("spgemm_long_rows", C.c_int32),
        ("spgemm_passes", C.c_int32),
        ("spgemm_esc_min", C.c_int32),
        ("bfs_group_cap", C.c_int32),
        ("sssp_hint", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


TUNE_BFS, TUNE_SSSP, TUNE_BFS_BATCH, TUNE_SSSP_BATCH = 1, 2, 4, 8


class Round(C.Structure):
    _fields_ = [("frontier", C.c_int32), ("pull", C.c_int32), ("edges", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [
        ("edges", C.c_int64),
        ("vertices", C.c_int64),
        ("improved", C.c_int64),
        ("iterations", C.c_int32),
        ("relax_rounds", C.c_int32),
        ("bytes_alg", C.c_double),
        ("kernel_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    """The loaded libapspie.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python apspie_amd/build.py` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    P, vp, i32, i64, u32 = C.POINTER, C.c_void_p, C.c_int32, C.c_int64, C.c_uint32
    sig = {
        "apspie_ctx_create": [C.c_int, P(vp)],
        "apspie_ctx_destroy": [vp],
        "apspie_ctx_stream": [vp],
        "apspie_ctx_synchronize": [vp],
        "apspie_ctx_cu_count": [vp],
        "apspie_ctx_set_profiling": [vp, C.c_int],
        "apspie_ctx_profile_read": [vp, C.c_int, P(C.c_double), P(i64)],
        "apspie_ctx_profile_reset": [vp],
        "apspie_hbm_copy_bandwidth": [vp, C.c_size_t, C.c_int, P(C.c_double)],
        "apspie_malloc": [vp, P(vp), C.c_size_t],
        "apspie_free": [vp, vp],
        "apspie_memcpy_h2d": [vp, vp, vp, C.c_size_t],
        "apspie_memcpy_d2h": [vp, vp, vp, C.c_size_t],
        "apspie_load_mtx": [C.c_char_p, u32, P(_Csr)],
        "apspie_csr_save": [C.c_char_p, P(_Csr)],
        "apspie_csr_attach_csc_device": [vp, P(_DCsr)],
        "apspie_coo_to_csr": [vp, vp, i64, i32, vp],
        "apspie_csr2csc": [vp, P(_DCsr), vp, vp, vp],
        "apspie_mxv": [vp, P(_DCsr), C.c_int, C.c_int, vp, vp, P(Stats)],
        "apspie_mxv_sparse": [vp, P(_DCsr), C.c_int, vp, vp, i32, vp, vp, P(i32), P(Stats)],
        "apspie_csr_load": [C.c_char_p, P(_Csr)],
        "apspie_host_coo_make_symmetric": [i64, vp, vp, vp, P(i64)],
        "apspie_host_coo_to_csr": [i32, i64, vp, vp, vp, vp, vp, vp],
        "apspie_rowptr_from_counts": [vp, vp, i32, vp, P(i64)],
        "apspie_host_csr_free": [P(_Csr)],
        "apspie_rmat": [C.c_int, C.c_int, u32, u32, u32, P(_Csr)],
        "apspie_ctx_set_tuning": [vp, P(Tuning)],
        "apspie_ctx_rounds": [vp, P(Round), C.c_int, P(C.c_int)],
        "apspie_ctx_get_tuning": [vp, P(Tuning)],
        "apspie_autotune": [vp, P(_DCsr), i32, i32, u32, P(Tuning)],
        "apspie_host_csr_transpose": [P(_Csr), P(_Csr)],
        "apspie_csr_upload": [vp, P(_Csr), P(_DCsr)],
        "apspie_csr_attach_csc": [vp, P(_DCsr), P(_Csr)],
        "apspie_csr_free": [vp, P(_DCsr)],
        "apspie_bfs": [vp, P(_DCsr), i32, vp, P(Stats)],
        "apspie_sssp": [vp, P(_DCsr), i32, vp, P(Stats)],
        "apspie_sssp_step": [vp, P(_DCsr), vp, vp, P(Stats)],
        "apspie_sssp_noalloc": [vp, P(_DCsr), i32, vp, P(Stats)],
        "apspie_mis": [vp, P(_DCsr), vp, vp, P(Stats)],
        "apspie_spmspv_or": [vp, P(_DCsr), vp, vp],
        "apspie_powerlaw_csr": [i32, i32, C.c_double, i32, u32, P(_Csr)],
        "apspie_spgemm": [vp, P(_DCsr), P(_DCsr), P(_DCsr), P(Stats)],
        "apspie_spgemm_outer": [vp, P(_DCsr), P(_DCsr), C.c_int32, P(_DCsr), P(Stats)],
        "apspie_strict_lower": [P(_Csr), P(_Csr)],
        "apspie_tc": [vp, P(_DCsr), P(i64), P(Stats)],
        "apspie_bfs_batch": [vp, P(_DCsr), vp, i32, vp, P(Stats)],
        "apspie_sssp_batch": [vp, P(_DCsr), vp, i32, vp, P(Stats)],
        "apspie_server_create": [C.c_int, P(_DCsr), C.c_int, vp, P(vp)],
        "apspie_server_submit": [vp, C.c_int, C.c_int, i32, vp, C.c_double, P(C.c_uint64)],
        "apspie_server_wait": [vp, C.c_uint64, P(C.c_double)],
        "apspie_server_drain": [vp],
        "apspie_server_stats": [vp, vp],
        "apspie_server_reset_stats": [vp],
        "apspie_server_set_tuning": [vp, P(Tuning)],
        "apspie_server_destroy": [vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.apspie_last_error.argtypes = [vp]
    L.apspie_last_error.restype = C.c_char_p
    L.apspie_ctx_stream.restype = vp
    L.apspie_host_csr_free.restype = None
    L.apspie_now.argtypes = []
    L.apspie_now.restype = C.c_double
    L.apspie_server_last_error.argtypes = [vp]
    L.apspie_server_last_error.restype = C.c_char_p
    _lib = L
    return L


def _check(rc: int, ctx=None, what: str = "") -> None:
    if rc != 0:
        msg = lib().apspie_last_error(ctx).decode() if ctx else what
        raise ApspieError(rc, f"{what}: {msg}" if ctx else msg)


# ---------------------------------------------------------------------------
# host graphs


@dataclass
class HostCsr:
    """Host CSR: int32 row_ptr[m+1], int32 col_ind[nnz], optional float32 val."""

    m: int
    n: int
    row_ptr: np.ndarray
    col_ind: np.ndarray
    val: Optional[np.ndarray] = None

    @property
    def nnz(self) -> int:
        return int(self.col_ind.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr)

    def argmax_degree(self) -> int:
        return int(np.argmax(self.degrees())) if self.m else 0

    def _as_c(self) -> _Csr:
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int32)
        self.col_ind = np.ascontiguousarray(self.col_ind, dtype=np.int32)
        if self.val is not None:
            self.val = np.ascontiguousarray(self.val, dtype=np.float32)
        s = _Csr()
        s.m, s.n, s.nnz = self.m, self.n, self.nnz
        s.row_ptr = self.row_ptr.ctypes.data
        s.col_ind = self.col_ind.ctypes.data if self.nnz else None
        s.val = self.val.ctypes.data if self.val is not None and self.nnz else None
        return s


def _host_from_c(s: _Csr) -> HostCsr:
    m, n, nnz = s.m, s.n, s.nnz
    rp = np.ctypeslib.as_array(C.cast(s.row_ptr, C.POINTER(C.c_int32)), shape=(m + 1,)).copy()
    ci = np.ctypeslib.as_array(C.cast(s.col_ind, C.POINTER(C.c_int32)), shape=(max(nnz, 1),))[:nnz].copy()
    val = None
    if s.val:
        val = np.ctypeslib.as_array(C.cast(s.val, C.POINTER(C.c_float)), shape=(max(nnz, 1),))[:nnz].copy()
    lib().apspie_host_csr_free(C.byref(s))
    return HostCsr(m, n, rp, ci, val)


def load_mtx(path: str, undirected: bool = False, sort_columns: bool = False, banner: bool = False) -> HostCsr:
    """readMtx + [makeSymmetric] + buildMatrix (src/util.cuh:355-643, src/matrix.cpp:55-111);
    banner: take field and symmetry from the %%MatrixMarket line."""
    s = _Csr()
    flags = ((LOAD_UNDIRECTED if undirected else 0) | (LOAD_SORT_COLUMNS if sort_columns else 0)
             | (LOAD_MM_BANNER if banner else 0))
    _check(lib().apspie_load_mtx(os.fsencode(path), flags, C.byref(s)), None, f"load_mtx({path})")
    return _host_from_c(s)


def save_csr(path: str, h: HostCsr) -> None:
    """Binary CSR cache file (apspie_csr_save)."""
    _check(lib().apspie_csr_save(os.fsencode(path), C.byref(h._as_c())), None, f"save_csr({path})")


def load_csr(path: str) -> HostCsr:
    """Read a binary CSR cache file (apspie_csr_load: check word and structure verified)."""
    s = _Csr()
    _check(lib().apspie_csr_load(os.fsencode(path), C.byref(s)), None, f"load_csr({path})")
    return _host_from_c(s)


def rmat(scale: int, edge_factor: int = 16, seed: int = 1, undirected: bool = False,
         weight_seed: Optional[int] = None) -> HostCsr:
    """Seeded RMAT(a=.57, b=.19, c=.19, d=.05) graph, deduplicated, sorted columns."""
    s = _Csr()
    flags = (RMAT_UNDIRECTED if undirected else 0) | (RMAT_WEIGHTED if weight_seed is not None else 0)
    _check(lib().apspie_rmat(scale, edge_factor, seed, flags, weight_seed or 0, C.byref(s)), None, "rmat")
    return _host_from_c(s)


def coo_make_symmetric(col: np.ndarray, row: np.ndarray, val: Optional[np.ndarray] = None):
    """makeSymmetric (src/util.cuh:572-643): returns (col, row, val) with the
    reversed non-loop entries added, sorted by (row, col), repeats dropped."""
    e = len(col)
    c = np.zeros(2 * e, np.int32)
    r = np.zeros(2 * e, np.int32)
    c[:e], r[:e] = col, row
    v = None
    if val is not None:
        v = np.zeros(2 * e, np.float32)
        v[:e] = val
    out = C.c_int64()
    _check(lib().apspie_host_coo_make_symmetric(e, c.ctypes.data, r.ctypes.data, v.ctypes.data if v is not None
                                                else None, C.byref(out)), None, "coo_make_symmetric")
    k = out.value
    return c[:k].copy(), r[:k].copy(), (v[:k].copy() if v is not None else None)


def coo_to_csr_host(m: int, row: np.ndarray, col: np.ndarray, val: Optional[np.ndarray] = None) -> HostCsr:
    """buildMatrix (src/matrix.cpp:55-111): stable counting sort by row."""
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nnz = len(row)
    rp = np.empty(m + 1, np.int32)
    ci = np.empty(max(nnz, 1), np.int32)
    va = np.empty(max(nnz, 1), np.float32) if val is not None else None
    cv = np.ascontiguousarray(val, dtype=np.float32) if val is not None else None
    _check(lib().apspie_host_coo_to_csr(m, nnz, row.ctypes.data, col.ctypes.data,
                                        cv.ctypes.data if cv is not None else None, rp.ctypes.data, ci.ctypes.data,
                                        va.ctypes.data if va is not None else None), None, "coo_to_csr_host")
    return HostCsr(m, m, rp, ci[:nnz], va[:nnz] if va is not None else None)


def transpose(h: HostCsr) -> HostCsr:
    """A^T by rows (A's CSC), entries of each row in increasing source-row order."""
    s = _Csr()
    hc = h._as_c()
    _check(lib().apspie_host_csr_transpose(C.byref(hc), C.byref(s)), None, "transpose")
    return _host_from_c(s)


def strict_lower(h: HostCsr) -> HostCsr:
    """buildLower: rows keep their columns j < i, sorted."""
    s = _Csr()
    hc = h._as_c()
    _check(lib().apspie_strict_lower(C.byref(hc), C.byref(s)), None, "strict_lower")
    return _host_from_c(s)


def powerlaw(m: int, n: int, mean_len: float = 8.0, max_len: int = 4096, seed: int = 3) -> HostCsr:
    """Seeded matrix with power-law row lengths, uniform sorted columns, U(-1,1) values."""
    s = _Csr()
    _check(lib().apspie_powerlaw_csr(m, n, mean_len, max_len, seed, C.byref(s)), None, "powerlaw")
    return _host_from_c(s)


# ---------------------------------------------------------------------------
# device side


class DeviceBuffer:
    """Device allocation owned by the library; freed with free() or on GC."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().apspie_malloc(ctx.h, C.byref(p), max(self.nbytes, 1)), ctx.h, "malloc")
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().apspie_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), self.ctx.h, "h2d")
        return self

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().apspie_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), self.ctx.h, "d2h")
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.h:
            lib().apspie_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceCsr:
    """CSR resident in HBM (library-owned), optionally with its CSC.

    csc: None (CSR only), "symmetric" (the CSR is its own transpose),
    "device" (built on the device, apspie_csr2csc), or the host transpose
    (transpose(h)) to upload beside it.
    """

    def __init__(self, ctx: "Context", h: HostCsr, csc=None):
        self.ctx = ctx
        self.m, self.n, self.nnz = h.m, h.n, h.nnz
        self.weighted = h.val is not None
        self._c = _DCsr()
        hc = h._as_c()
        _check(lib().apspie_csr_upload(ctx.h, C.byref(hc), C.byref(self._c)), ctx.h, "csr_upload")
        if csc is not None:
            if isinstance(csc, str):
                assert csc in ("symmetric", "device"), csc
                if csc == "device":
                    rc = lib().apspie_csr_attach_csc_device(ctx.h, C.byref(self._c))
                else:
                    rc = lib().apspie_csr_attach_csc(ctx.h, C.byref(self._c), None)
            else:
                tc_ = csc._as_c()
                rc = lib().apspie_csr_attach_csc(ctx.h, C.byref(self._c), C.byref(tc_))
            if rc != 0:
                self.free()
                _check(rc, ctx.h, "csr_attach_csc")

    @property
    def has_csc(self) -> bool:
        return bool(self._c.csc_col_ptr)

    @property
    def c(self) -> _DCsr:
        return self._c

    def free(self) -> None:
        if self._c.row_ptr and self.ctx.h:
            lib().apspie_csr_free(self.ctx.h, C.byref(self._c))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """A device, a HIP stream and a scratch pool (apspie_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().apspie_ctx_create(device, C.byref(h)), None, "apspie_ctx_create")
        self.h = h.value

    def close(self) -> None:
        if self.h:
            lib().apspie_ctx_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def cu_count(self) -> int:
        return lib().apspie_ctx_cu_count(self.h)

    def synchronize(self) -> None:
        _check(lib().apspie_ctx_synchronize(self.h), self.h, "synchronize")

    # -- per-kernel-class event timing ---------------------------------------
    KCLASS = {"expand": 0, "compact": 1, "other": 2, "pull": 3, "bin": 4, "apply": 5, "round": 6}

    def set_profiling(self, on: bool) -> None:
        _check(lib().apspie_ctx_set_profiling(self.h, int(on)), self.h, "set_profiling")

    def profile_reset(self) -> None:
        _check(lib().apspie_ctx_profile_reset(self.h), self.h, "profile_reset")

    def profile(self) -> dict:
        """{kernel class: (device ms, launches)} since the last reset."""
        out = {}
        for name, k in self.KCLASS.items():
            ms, n = C.c_double(), C.c_int64()
            _check(lib().apspie_ctx_profile_read(self.h, k, C.byref(ms), C.byref(n)), self.h, "profile_read")
            out[name] = (ms.value, n.value)
        return out

    def hbm_copy_bandwidth(self, nbytes: int = 1 << 30, iters: int = 5) -> float:
        """Best device-copy rate in GB/s (read + write bytes)."""
        g = C.c_double()
        _check(lib().apspie_hbm_copy_bandwidth(self.h, nbytes, iters, C.byref(g)), self.h, "copy")
        return g.value

    def upload(self, h: HostCsr, csc=None) -> DeviceCsr:
        return DeviceCsr(self, h, csc)

    def rounds(self) -> list:
        """Per-round frontier log of the last traversal: [(vertices, edges, pull)]."""
        buf = (Round * 64)()
        n = C.c_int()
        _check(lib().apspie_ctx_rounds(self.h, buf, 64, C.byref(n)), self.h, "rounds")
        return [(buf[i].frontier, buf[i].edges, buf[i].pull) for i in range(n.value)]

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().apspie_ctx_get_tuning(self.h, C.byref(t)), self.h, "get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        cur = self.tuning()
        unknown = set(kw) - set(cur)
        if unknown:
            raise KeyError(f"unknown tuning knobs {sorted(unknown)}")
        cur.update(kw)
        t = Tuning(**cur)
        _check(lib().apspie_ctx_set_tuning(self.h, C.byref(t)), self.h, "set_tuning")

    def autotune(self, g: "DeviceCsr", src: int, kinds: int = TUNE_BFS | TUNE_SSSP, k: int = 64) -> dict:
        """Time candidate knob values on g (apspie_autotune), keep the fastest
        in this context and return them."""
        t = Tuning()
        _check(lib().apspie_autotune(self.h, C.byref(g.c), src, k, kinds, C.byref(t)), self.h, "autotune")
        return t.as_dict()

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    # -- plain path --------------------------------------------------------
    def bfs_into(self, g: DeviceCsr, src: int, d_depth: DeviceBuffer) -> Stats:
        st = Stats()
        _check(lib().apspie_bfs(self.h, C.byref(g.c), src, d_depth.ptr, C.byref(st)), self.h, "bfs")
        return st

    def bfs(self, g: DeviceCsr, src: int):
        """Returns (depth int32[m], Stats); -1 marks unreached vertices."""
        d = self.alloc(4 * max(g.m, 1))
        st = self.bfs_into(g, src, d)
        out = d.download(np.int32, g.m)
        d.free()
        return out, st

    def sssp_into(self, g: DeviceCsr, src: int, d_dist: DeviceBuffer) -> Stats:
        st = Stats()
        _check(lib().apspie_sssp(self.h, C.byref(g.c), src, d_dist.ptr, C.byref(st)), self.h, "sssp")
        return st

    def sssp_noalloc_into(self, g: DeviceCsr, src: int, d_dist: DeviceBuffer) -> Stats:
        """apspie_sssp_noalloc: the same traversal with its scratch in the
        library's fixed device array (no allocation in the call)."""
        st = Stats()
        _check(lib().apspie_sssp_noalloc(self.h, C.byref(g.c), src, d_dist.ptr, C.byref(st)), self.h,
               "sssp_noalloc")
        return st

    def sssp(self, g: DeviceCsr, src: int):
        """Returns (dist float32[m], Stats); SSSP_INF marks unreachable vertices."""
        d = self.alloc(4 * max(g.m, 1))
        st = self.sssp_into(g, src, d)
        out = d.download(np.float32, g.m)
        d.free()
        return out, st

    def sssp_step(self, g: DeviceCsr, x: np.ndarray):
        """One min-plus SpMSpV step y = min-plus(A^T, x); returns (y, Stats)."""
        dx = self.alloc(4 * max(g.m, 1)).upload(np.asarray(x, dtype=np.float32))
        dy = self.alloc(4 * max(g.m, 1))
        st = Stats()
        _check(lib().apspie_sssp_step(self.h, C.byref(g.c), dx.ptr, dy.ptr, C.byref(st)), self.h, "sssp_step")
        y = dy.download(np.float32, g.m)
        dx.free()
        dy.free()
        return y, st

    def mis(self, g: DeviceCsr, prio: np.ndarray):
        """Greedy-by-(prio, -id) MIS of a symmetric graph; returns (int32[m] 0/1, Stats)."""
        dp = self.alloc(4 * max(g.m, 1)).upload(np.asarray(prio, dtype=np.uint32))
        dm = self.alloc(4 * max(g.m, 1))
        st = Stats()
        _check(lib().apspie_mis(self.h, C.byref(g.c), dp.ptr, dm.ptr, C.byref(st)), self.h, "mis")
        out = dm.download(np.int32, g.m)
        dp.free()
        dm.free()
        return out, st

    def spgemm_device(self, a: DeviceCsr, b: DeviceCsr):
        """C = A * B left on the device: returns (_DCsr owned by the library, Stats)."""
        c = _DCsr()
        st = Stats()
        _check(lib().apspie_spgemm(self.h, C.byref(a.c), C.byref(b.c), C.byref(c), C.byref(st)), self.h, "spgemm")
        return c, st

    def free_owned(self, c: _DCsr) -> None:
        lib().apspie_csr_free(self.h, C.byref(c))

    def download_csr(self, c: _DCsr) -> HostCsr:
        m, nnz = c.m, c.nnz
        rp = np.empty(m + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        va = np.empty(nnz, np.float32)
        for host, dev in ((rp, c.row_ptr), (ci, c.col_ind), (va, c.val)):
            if host.nbytes:
                _check(lib().apspie_memcpy_d2h(self.h, host.ctypes.data, dev, host.nbytes), self.h, "d2h")
        return HostCsr(m, c.n, rp, ci, va)

    def spgemm(self, a: DeviceCsr, b: DeviceCsr):
        """C = A * B; returns (HostCsr C, Stats)."""
        c, st = self.spgemm_device(a, b)
        try:
            return self.download_csr(c), st
        finally:
            self.free_owned(c)

    def spgemm_outer_device(self, a: DeviceCsr, b: DeviceCsr, part_size: int = 65536):
        """C = A * B^T by the strip-pair outer products (apspie_spgemm_outer),
        left on the device: returns (_DCsr owned by the library, Stats)."""
        c = _DCsr()
        st = Stats()
        _check(lib().apspie_spgemm_outer(self.h, C.byref(a.c), C.byref(b.c), int(part_size), C.byref(c),
                                         C.byref(st)), self.h, "spgemm_outer")
        return c, st

    def spgemm_outer(self, a: DeviceCsr, b: DeviceCsr, part_size: int = 65536):
        """C = A * B^T by the strip-pair outer products; returns (HostCsr C, Stats)."""
        c, st = self.spgemm_outer_device(a, b, part_size)
        try:
            return self.download_csr(c), st
        finally:
            self.free_owned(c)

    def bfs_batch(self, g: DeviceCsr, srcs):
        """BFS from up to 256 sources; returns (int32[k, m] depths, Stats)."""
        srcs = np.ascontiguousarray(srcs, dtype=np.int32)
        k = len(srcs)
        ds = self.alloc(4 * max(k, 1)).upload(srcs)
        dd = self.alloc(4 * max(k * g.m, 1))
        st = Stats()
        _check(lib().apspie_bfs_batch(self.h, C.byref(g.c), ds.ptr, k, dd.ptr, C.byref(st)), self.h, "bfs_batch")
        out = dd.download(np.int32, k * g.m).reshape(k, g.m)
        ds.free()
        dd.free()
        return out, st

    def sssp_batch(self, g: DeviceCsr, srcs):
        """SSSP from up to 64 sources; returns (float32[k, m] distances, Stats)."""
        srcs = np.ascontiguousarray(srcs, dtype=np.int32)
        k = len(srcs)
        ds = self.alloc(4 * max(k, 1)).upload(srcs)
        dd = self.alloc(4 * max(k * g.m, 1))
        st = Stats()
        _check(lib().apspie_sssp_batch(self.h, C.byref(g.c), ds.ptr, k, dd.ptr, C.byref(st)), self.h, "sssp_batch")
        out = dd.download(np.float32, k * g.m).reshape(k, g.m)
        ds.free()
        dd.free()
        return out, st

    def mxv(self, g: DeviceCsr, x: np.ndarray, semiring: int = SR_PLUS_TIMES, above_half: bool = False):
        """apspie_mxv: dense x[m] -> dense y[n] over the semiring (present: x not the
        identity, or x > 0.5 when above_half); returns (y, Stats)."""
        dx = self.alloc(4 * max(g.m, 1)).upload(np.asarray(x, dtype=np.float32))
        dy = self.alloc(4 * max(g.n, 1))
        st = Stats()
        _check(lib().apspie_mxv(self.h, C.byref(g.c), semiring, 1 if above_half else 0, dx.ptr, dy.ptr,
                                C.byref(st)), self.h, "mxv")
        y = dy.download(np.float32, g.n)
        dx.free()
        dy.free()
        return y, st

    def mxv_sparse(self, g: DeviceCsr, ind: np.ndarray, val: np.ndarray, semiring: int = SR_PLUS_TIMES):
        """apspie_mxv_sparse: (index, value) pairs -> (reached columns ascending, their values)."""
        ind = np.ascontiguousarray(ind, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float32)
        nx = len(ind)
        di = self.alloc(4 * max(nx, 1)).upload(ind) if nx else None
        dv = self.alloc(4 * max(nx, 1)).upload(val) if nx else None
        yi = self.alloc(4 * max(g.n, 1))
        yv = self.alloc(4 * max(g.n, 1))
        ny = C.c_int32()
        st = Stats()
        _check(lib().apspie_mxv_sparse(self.h, C.byref(g.c), semiring, di.ptr if di else None,
                                       dv.ptr if dv else None, nx, yi.ptr, yv.ptr, C.byref(ny), C.byref(st)),
               self.h, "mxv_sparse")
        out = yi.download(np.int32, ny.value), yv.download(np.float32, ny.value)
        for b in (di, dv, yi, yv):
            if b:
                b.free()
        return out

    def csr2csc(self, g: DeviceCsr, values: bool = True) -> HostCsr:
        """apspie_csr2csc on the device; returns the CSC as a host CSR of A^T."""
        cp = self.alloc(4 * (g.n + 1))
        ri = self.alloc(4 * max(g.nnz, 1))
        cv = self.alloc(4 * max(g.nnz, 1)) if (values and g.weighted) else None
        _check(lib().apspie_csr2csc(self.h, C.byref(g.c), cp.ptr, ri.ptr, cv.ptr if cv else None), self.h, "csr2csc")
        out = HostCsr(g.n, g.m, cp.download(np.int32, g.n + 1), ri.download(np.int32, g.nnz),
                      cv.download(np.float32, g.nnz) if cv else None)
        for b in (cp, ri, cv):
            if b:
                b.free()
        return out

    def coo_to_csr(self, rows: np.ndarray, m: int) -> np.ndarray:
        """apspie_coo_to_csr: row offsets of sorted COO row indices."""
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        dr = self.alloc(4 * max(len(rows), 1)).upload(rows) if len(rows) else None
        rp = self.alloc(4 * (m + 1))
        _check(lib().apspie_coo_to_csr(self.h, dr.ptr if dr else None, len(rows), m, rp.ptr), self.h, "coo_to_csr")
        out = rp.download(np.int32, m + 1)
        rp.free()
        if dr:
            dr.free()
        return out

    def rowptr_from_counts(self, counts: np.ndarray):
        """apspie_rowptr_from_counts: (row_ptr int32[n+1], exact total)."""
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        n = len(counts)
        dc = self.alloc(4 * max(n, 1)).upload(counts) if n else None
        rp = self.alloc(4 * (n + 1))
        tot = C.c_int64()
        try:
            _check(lib().apspie_rowptr_from_counts(self.h, dc.ptr if dc else None, n, rp.ptr, C.byref(tot)),
                   self.h, "rowptr_from_counts")
            return rp.download(np.int32, n + 1), tot.value
        finally:
            rp.free()
            if dc:
                dc.free()

    def tc(self, l: DeviceCsr):
        """Triangles from the strict lower triangle L; returns (int count, Stats)."""
        cnt = C.c_int64()
        st = Stats()
        _check(lib().apspie_tc(self.h, C.byref(l.c), C.byref(cnt), C.byref(st)), self.h, "tc")
        return int(cnt.value), st

    def spmspv_or(self, g: DeviceCsr, x: np.ndarray) -> np.ndarray:
        """Boolean SpMSpV: out[v] = 1 iff some u with x[u] != 0 has v in its row."""
        dx = self.alloc(4 * max(g.m, 1)).upload(np.asarray(x, dtype=np.int32))
        dy = self.alloc(4 * max(g.m, 1))
        _check(lib().apspie_spmspv_or(self.h, C.byref(g.c), dx.ptr, dy.ptr), self.h, "spmspv_or")
        y = dy.download(np.int32, g.m)
        dx.free()
        dy.free()
        return y


class ServerOpts(C.Structure):
    _fields_ = [
        ("lanes", C.c_int),
        ("max_bfs_batch", C.c_int),
        ("max_sssp_batch", C.c_int),
        ("plain_bfs_upto", C.c_int),
        ("plain_sssp_upto", C.c_int),
        ("budget_bytes", C.c_size_t),
    ]


class ServerReport(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("completed", C.c_int64),
        ("rejected", C.c_int64),
        ("batches", C.c_int64),
        ("mean_batch", C.c_double),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("max_ms", C.c_double),
        ("qps", C.c_double),
        ("trivial", C.c_int64),
        ("joined", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_BFS, REQ_SSSP = 0, 1


def now() -> float:
    """The serving loop's steady clock, in seconds."""
    return lib().apspie_now()


class Server:
    """apspie_server: batched BFS/SSSP service over device-resident graphs."""

    def __init__(self, graphs: Sequence[DeviceCsr], lanes: int = 4, budget_bytes: int = 32 << 30,
                 max_bfs_batch: int = 256, max_sssp_batch: int = 64, device: int = 0,
                 plain_bfs_upto: int = 0, plain_sssp_upto: int = 0):
        """plain_*_upto: batches of at most that many requests run as plain
        launches (0: the library default, -1: never)."""
        L = lib()
        self._graphs = list(graphs)  # keep the device CSRs alive
        self._held = {}  # ticket -> output array the server may still write
        arr = (_DCsr * len(graphs))(*[g.c for g in graphs])
        opts = ServerOpts(lanes, max_bfs_batch, max_sssp_batch, plain_bfs_upto, plain_sssp_upto, budget_bytes)
        h = C.c_void_p()
        _check(L.apspie_server_create(device, arr, len(graphs), C.byref(opts), C.byref(h)), None, "server_create")
        self.h = h.value

    def submit(self, kind: int, graph: int, src: int, out: Optional[np.ndarray] = None,
               arrival: float = -1.0) -> int:
        """Queue one request; `out` (or None) receives the row: a C-contiguous,
        writeable int32 (BFS) or float32 (SSSP) array of at least m entries.
        The server writes it from its own threads after this call returns, so
        the array is held here until wait(), drain() or close() releases it."""
        if out is not None:
            want = np.int32 if kind == REQ_BFS else np.float32
            if not isinstance(out, np.ndarray) or out.dtype != want:
                raise TypeError(f"out must be a numpy {np.dtype(want).name} array")
            if not (out.flags.c_contiguous and out.flags.writeable):
                raise ValueError("out must be C-contiguous and writeable")
            if not 0 <= graph < len(self._graphs) or out.size < self._graphs[graph].m:
                raise ValueError(f"out holds {out.size} entries, the graph has "
                                 f"{self._graphs[graph].m if 0 <= graph < len(self._graphs) else '?'} vertices")
        t = C.c_uint64()
        ptr = out.ctypes.data if out is not None else None
        rc = lib().apspie_server_submit(self.h, kind, graph, src, ptr, arrival, C.byref(t))
        if rc != 0:
            raise ApspieError(rc, "server_submit")
        if out is not None:
            self._held[t.value] = out
        return t.value

    def wait(self, ticket: int) -> float:
        lat = C.c_double()
        rc = lib().apspie_server_wait(self.h, ticket, C.byref(lat))
        self._held.pop(ticket, None)
        if rc != 0:
            raise ApspieError(rc, lib().apspie_server_last_error(self.h).decode())
        return lat.value

    def drain(self) -> None:
        lib().apspie_server_drain(self.h)
        self._held.clear()

    def report(self) -> dict:
        r = ServerReport()
        lib().apspie_server_stats(self.h, C.byref(r))
        return r.as_dict()

    def reset_stats(self) -> None:
        lib().apspie_server_reset_stats(self.h)

    def set_tuning(self, tuning: dict) -> None:
        """Knobs of every lane's context: every knob, as Context.tuning()
        returns them (apspie_server_set_tuning)."""
        names = {k for k, _ in Tuning._fields_}
        if set(tuning) != names:
            raise KeyError(f"tuning must name exactly the knobs {sorted(names)}")
        t = Tuning(**{k: int(v) for k, v in tuning.items()})
        rc = lib().apspie_server_set_tuning(self.h, C.byref(t))
        if rc != 0:
            raise ApspieError(rc, "server_set_tuning: " + lib().apspie_server_last_error(self.h).decode())

    def close(self) -> None:
        if self.h:
            lib().apspie_server_destroy(self.h)  # joins the workers: no write is left in flight
        self.h = None
        self._held.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def exported_symbols(header: str) -> Sequence[str]:
    """Function names declared in a C header (for ABI completeness checks)."""
    import re

    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(apspie_[a-z0-9_]+)\s*\(", text)))
