"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_MT_PATH = os.path.join(HERE, "_ref", "libref_mt.so")


class OrcCsr(C.Structure):
    _fields_ = [
        ("m", C.c_int32),
        ("n", C.c_int32),
        ("nnz", C.c_int64),
        ("row_ptr", C.POINTER(C.c_int32)),
        ("col_ind", C.POINTER(C.c_int32)),
        ("val", C.POINTER(C.c_float)),
    ]


@dataclass
class Csr:
    """Host CSR as numpy arrays (int32 row_ptr/col_ind, optional float32 val)."""

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


def build(ref: bool = False) -> None:
    """Compile the oracle (and oracle/_ref when the reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_mt_fill.argtypes = [C.c_uint32, C.c_int64, P(C.c_uint32)]
        L.orc_csr_free.argtypes = [P(OrcCsr)]
        L.orc_read_mtx.argtypes = [C.c_char_p, C.c_int, P(OrcCsr)]
        L.orc_rmat.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_int, P(OrcCsr)]
        L.orc_rmat_weights.argtypes = [P(OrcCsr), C.c_uint32]
        L.orc_argmax_degree.argtypes = [P(OrcCsr)]
        L.orc_argmax_degree.restype = C.c_int32
        L.orc_bfs.argtypes = [P(OrcCsr), C.c_int32, C.c_int32, P(C.c_int32), P(C.c_int64), P(C.c_int64)]
        L.orc_bfs.restype = C.c_int32
        L.orc_sssp_dijkstra.argtypes = [P(OrcCsr), C.c_int32, P(C.c_float)]
        L.orc_sssp_bf.argtypes = [P(OrcCsr), C.c_int32, P(C.c_float), P(C.c_int64), P(C.c_int64), P(C.c_int64)]
        L.orc_sssp_bf.restype = C.c_int32
        L.orc_mis_greedy.argtypes = [P(OrcCsr), P(C.c_uint32), P(C.c_int32)]
        L.orc_mis_index.argtypes = [P(OrcCsr), C.c_int32, P(C.c_int32)]
        L.orc_spmspv_or.argtypes = [P(OrcCsr), P(C.c_int32), P(C.c_int32)]
        L.orc_sssp_step.argtypes = [P(OrcCsr), P(C.c_float), P(C.c_float)]
        L.orc_spgemm.argtypes = [P(OrcCsr), P(OrcCsr), P(OrcCsr), P(C.c_int64)]
        L.orc_strict_lower.argtypes = [P(OrcCsr), P(OrcCsr)]
        L.orc_tc.argtypes = [P(OrcCsr), P(C.c_int64)]
        L.orc_tc.restype = C.c_int64
        L.orc_tc_rows.argtypes = [P(OrcCsr), C.c_int32, C.c_int32, P(C.c_int64)]
        L.orc_tc_rows.restype = C.c_int64
        L.orc_mxv.argtypes = [P(OrcCsr), C.c_int, P(C.c_float), P(C.c_uint8), P(C.c_float), P(C.c_uint8)]
        L.orc_mxv.restype = None
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray], ct):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    return a.ctypes.data_as(C.POINTER(ct))


def _to_c(a: Csr):
    """Borrow numpy arrays as an OrcCsr (keep `a` alive while in use)."""
    a.row_ptr = np.ascontiguousarray(a.row_ptr, dtype=np.int32)
    a.col_ind = np.ascontiguousarray(a.col_ind, dtype=np.int32)
    if a.val is not None:
        a.val = np.ascontiguousarray(a.val, dtype=np.float32)
    s = OrcCsr()
    s.m, s.n, s.nnz = a.m, a.n, a.nnz
    s.row_ptr = _ptr(a.row_ptr, C.c_int32)
    s.col_ind = _ptr(a.col_ind, C.c_int32)
    s.val = _ptr(a.val, C.c_float)
    return s


def _from_c(s: OrcCsr) -> Csr:
    m, nnz = s.m, s.nnz
    rp = np.ctypeslib.as_array(s.row_ptr, shape=(m + 1,)).copy()
    ci = np.ctypeslib.as_array(s.col_ind, shape=(max(nnz, 1),))[:nnz].copy()
    val = None
    if bool(s.val):
        val = np.ctypeslib.as_array(s.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib().orc_csr_free(C.byref(s))
    return Csr(m, s.n, rp, ci, val)


def mt_fill(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt_fill(seed, n, _ptr(out, C.c_uint32))
    return out


def ref_mt_fill(seed: int, n: int) -> Optional[np.ndarray]:
    """The reference's own genrand_int32 stream, when oracle/_ref was built."""
    if not os.path.exists(REF_MT_PATH):
        return None
    L = C.CDLL(REF_MT_PATH)
    L.ref_mt_fill.argtypes = [C.c_uint32, C.c_int64, C.POINTER(C.c_uint32)]
    out = np.empty(n, dtype=np.uint32)
    L.ref_mt_fill(seed, n, _ptr(out, C.c_uint32))
    return out


def read_mtx(path: str, undirected: bool = False) -> Csr:
    s = OrcCsr()
    rc = lib().orc_read_mtx(path.encode(), int(undirected), C.byref(s))
    if rc != 0:
        raise ValueError(f"orc_read_mtx({path}) failed: {rc}")
    return _from_c(s)


def rmat(scale: int, edge_factor: int, seed: int, undirected: bool, weight_seed: Optional[int] = None) -> Csr:
    s = OrcCsr()
    rc = lib().orc_rmat(scale, edge_factor, seed, int(undirected), C.byref(s))
    if rc != 0:
        raise ValueError(f"orc_rmat failed: {rc}")
    if weight_seed is not None:
        lib().orc_rmat_weights(C.byref(s), weight_seed)
    return _from_c(s)


def argmax_degree(a: Csr) -> int:
    s = _to_c(a)
    return int(lib().orc_argmax_degree(C.byref(s)))


def bfs(a: Csr, src: int, stop: int = 1 << 30):
    """Returns (depth int32[m], search_depth, e_exp, v_reached)."""
    s = _to_c(a)
    depth = np.empty(a.m, dtype=np.int32)
    e, v = C.c_int64(), C.c_int64()
    sd = lib().orc_bfs(C.byref(s), src, stop, _ptr(depth, C.c_int32), C.byref(e), C.byref(v))
    return depth, int(sd), int(e.value), int(v.value)


def sssp_dijkstra(a: Csr, src: int) -> np.ndarray:
    s = _to_c(a)
    dist = np.empty(a.m, dtype=np.float32)
    lib().orc_sssp_dijkstra(C.byref(s), src, _ptr(dist, C.c_float))
    return dist


def sssp_bf(a: Csr, src: int):
    """Returns (dist, iterations, e_relax, sum_f, r_imp); -1 = unreachable."""
    s = _to_c(a)
    dist = np.empty(a.m, dtype=np.float32)
    er, sf, ri = C.c_int64(), C.c_int64(), C.c_int64()
    it = lib().orc_sssp_bf(C.byref(s), src, _ptr(dist, C.c_float), C.byref(er), C.byref(sf), C.byref(ri))
    return dist, int(it), int(er.value), int(sf.value), int(ri.value)


def sssp_step(a: Csr, x: np.ndarray) -> np.ndarray:
    s = _to_c(a)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(a.m, dtype=np.float32)
    lib().orc_sssp_step(C.byref(s), _ptr(x, C.c_float), _ptr(y, C.c_float))
    return y


def mis_greedy(a: Csr, prio: np.ndarray) -> np.ndarray:
    s = _to_c(a)
    prio = np.ascontiguousarray(prio, dtype=np.uint32)
    out = np.empty(a.m, dtype=np.int32)
    lib().orc_mis_greedy(C.byref(s), _ptr(prio, C.c_uint32), _ptr(out, C.c_int32))
    return out


def mis_index(a: Csr, src: int = 0) -> np.ndarray:
    s = _to_c(a)
    out = np.empty(a.m, dtype=np.int32)
    lib().orc_mis_index(C.byref(s), src, _ptr(out, C.c_int32))
    return out


def spmspv_or(a: Csr, x: np.ndarray) -> np.ndarray:
    s = _to_c(a)
    x = np.ascontiguousarray(x, dtype=np.int32)
    out = np.empty(a.m, dtype=np.int32)
    lib().orc_spmspv_or(C.byref(s), _ptr(x, C.c_int32), _ptr(out, C.c_int32))
    return out


def spgemm(a: Csr, b: Csr):
    """Returns (C, products)."""
    sa, sb, sc = _to_c(a), _to_c(b), OrcCsr()
    p = C.c_int64()
    rc = lib().orc_spgemm(C.byref(sa), C.byref(sb), C.byref(sc), C.byref(p))
    if rc != 0:
        raise ValueError(f"orc_spgemm failed: {rc}")
    return _from_c(sc), int(p.value)


def strict_lower(a: Csr) -> Csr:
    sa, sl = _to_c(a), OrcCsr()
    lib().orc_strict_lower(C.byref(sa), C.byref(sl))
    return _from_c(sl)


def tc(l: Csr):
    """Returns (triangles, wedges)."""
    s = _to_c(l)
    w = C.c_int64()
    t = lib().orc_tc(C.byref(s), C.byref(w))
    return int(t), int(w.value)


def tc_rows(l: Csr, r0: int, r1: int):
    """The part of tc() whose mask rows are r0..r1-1: (triangles, wedges)."""
    s = _to_c(l)
    w = C.c_int64()
    t = lib().orc_tc_rows(C.byref(s), r0, r1, C.byref(w))
    return int(t), int(w.value)


def mxv(a: Csr, sr: int, x: np.ndarray, listed: Optional[np.ndarray] = None):
    """General-semiring SpMSpV: returns (y float32[n], reached uint8[n])."""
    s = _to_c(a)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty(a.n, dtype=np.float32)
    r = np.empty(a.n, dtype=np.uint8)
    ls = None if listed is None else np.ascontiguousarray(listed, dtype=np.uint8)
    lib().orc_mxv(C.byref(s), sr, _ptr(x, C.c_float), _ptr(ls, C.c_uint8), _ptr(y, C.c_float), _ptr(r, C.c_uint8))
    return y, r
