"""CPU oracle for the boolean SpGEMM path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings over ``oracle/liboracle.so`` (a plain-C restatement of the
clBool host functions, built by ``oracle/Makefile``) plus the reference's own
``FastRandom`` built into ``oracle/_ref``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import
this package, and only as the checker or the timed CPU baseline.  The product
library (``clbool_amd``) never imports it.

Each wrapper names the reference file:line it restates (paths are relative to
the clBool source tree); the C file repeats the citations line by line.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class _OrcMat(C.Structure):
    _fields_ = [("nzr", C.c_uint32), ("nnz", C.c_uint64), ("rpt", u32p), ("rows", u32p), ("cols", u32p)]


class _OrcPairs(C.Structure):
    _fields_ = [("n", C.c_uint64), ("rows", u32p), ("cols", u32p)]


class _Dims(C.Structure):
    _fields_ = [("F", C.c_uint32), ("H1", C.c_uint32), ("H2", C.c_uint32)]


def lib():
    """Load liboracle.so (building it with make when it is missing)."""
    global _LIB
    if _LIB is None:
        # CLBOOL_ORACLE_LIB: another build of the same source (the sanitizer
        # build of the Makefile's `sanitize` target)
        path = os.environ.get("CLBOOL_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_generate_coo_pairs.restype = C.POINTER(_OrcPairs)
        L.orc_generate_coo_pairs.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_generate_coo.restype = C.POINTER(_OrcPairs)
        L.orc_generate_coo.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_pairs_to_dcsr.restype = C.POINTER(_OrcMat)
        L.orc_pairs_to_dcsr.argtypes = [u32p, u32p, C.c_uint64]
        L.orc_spgemm_reference.restype = C.POINTER(_OrcMat)
        L.orc_spgemm_reference.argtypes = [u32p, u32p, u32p, C.c_uint32, u32p, u32p, u32p, C.c_uint32]
        L.orc_spgemm_gustavson.restype = C.POINTER(_OrcMat)
        L.orc_spgemm_gustavson.argtypes = [u32p, u32p, u32p, C.c_uint32, u32p, u32p, u32p, C.c_uint32,
                                           C.c_uint32, C.c_uint32, u32p, C.c_uint32, C.c_int]
        L.orc_spgemm_digest.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, u32p, u32p, u32p, C.c_uint32,
                                        C.c_uint32, C.c_int, u64p, u64p]
        L.orc_dcsr_row_digest.restype = C.c_uint64
        L.orc_dcsr_row_digest.argtypes = [u32p, u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, u64p]
        L.orc_dense_bool_mm.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_workload.argtypes = [u32p, u32p, C.c_uint32, u32p, u32p, C.c_uint32, u64p]
        L.orc_coo_add.restype = C.POINTER(_OrcPairs)
        L.orc_coo_add.argtypes = [u32p, u32p, C.c_uint64, u32p, u32p, C.c_uint64]
        L.orc_kron.restype = C.POINTER(_OrcPairs)
        L.orc_kron.argtypes = [u32p, u32p, C.c_uint64, u32p, u32p, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_transpose.restype = C.POINTER(_OrcMat)
        L.orc_transpose.argtypes = [u32p, u32p, u32p, C.c_uint32]
        L.orc_submatrix.restype = C.POINTER(_OrcMat)
        L.orc_submatrix.argtypes = [u32p, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_reduce.restype = C.POINTER(_OrcMat)
        L.orc_reduce.argtypes = [u32p, C.c_uint32]
        L.orc_exclusive_scan_u32.restype = C.c_uint64
        L.orc_exclusive_scan_u32.argtypes = [u32p, C.c_uint64]
        L.orc_csr_from_pairs.restype = C.POINTER(_OrcMat)
        L.orc_csr_from_pairs.argtypes = [u32p, u32p, C.c_uint64, C.c_uint32]
        L.orc_dcsr_to_csr_rpt.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, u32p]
        L.orc_csr_to_dcsr.restype = C.POINTER(_OrcMat)
        L.orc_csr_to_dcsr.argtypes = [u32p, C.c_uint32]
        L.orc_coo_build.restype = C.POINTER(_OrcPairs)
        L.orc_coo_build.argtypes = [u32p, u32p, C.c_uint64, C.c_int, C.c_int]
        L.orc_fr_stream.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_uint32, C.POINTER(C.c_int32)]
        L.orc_free.argtypes = [C.POINTER(_OrcMat)]
        L.orc_free_pairs.argtypes = [C.POINTER(_OrcPairs)]
        L.orc_mlp_param_count.restype = C.c_size_t
        L.orc_mlp_param_count.argtypes = [_Dims]
        L.orc_mlp_forward.argtypes = [_Dims, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = [_Dims, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_adam_step.argtypes = [C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own FastRandom (oracle/_ref), or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libfastrandom_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_fastrandom_stream.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_uint32, C.POINTER(C.c_int32)]
        L.ref_unsigned_long_bits.restype = C.c_int32
        _REF = L
    return _REF


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _p32(a):
    return a.ctypes.data_as(u32p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _take_pairs(ptr):
    p = ptr.contents
    n = int(p.n)
    rows = np.ctypeslib.as_array(p.rows, shape=(max(n, 1),))[:n].copy()
    cols = np.ctypeslib.as_array(p.cols, shape=(max(n, 1),))[:n].copy()
    lib().orc_free_pairs(ptr)
    return rows, cols


@dataclass
class Dcsr:
    """Host DCSR matrix (matrix_dcsr_cpu, src/core/cpu_matrices.hpp:9-51)."""
    rpt: np.ndarray
    rows: np.ndarray
    cols: np.ndarray
    nrows: int = 0
    ncols: int = 0

    @property
    def nnz(self):
        return int(self.cols.size)

    @property
    def nzr(self):
        return int(self.rows.size)


def _take_mat(ptr, nrows=0, ncols=0, with_rows=True):
    m = ptr.contents
    nzr, nnz = int(m.nzr), int(m.nnz)
    rpt = np.ctypeslib.as_array(m.rpt, shape=(nzr + 1,)).copy() if m.rpt else np.zeros(1, np.uint32)
    rows = (np.ctypeslib.as_array(m.rows, shape=(max(nzr, 1),))[:nzr].copy()
            if (with_rows and m.rows) else np.zeros(0, np.uint32))
    cols = (np.ctypeslib.as_array(m.cols, shape=(max(nnz, 1),))[:nnz].copy()
            if m.cols else np.zeros(0, np.uint32))
    lib().orc_free(ptr)
    return Dcsr(rpt, rows, cols, nrows, ncols)


# --------------------------------------------------------------------------
def fastrandom_stream(seed, lo, hi, n):
    """FastRandom draws (fast_random.h:22-32), oracle restatement."""
    out = np.zeros(n, np.int32)
    lib().orc_fr_stream(seed, lo, hi, n, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def ref_fastrandom_stream(seed, lo, hi, n):
    """FastRandom draws from the reference's own header (oracle/_ref)."""
    L = ref_lib()
    if L is None:
        return None
    out = np.zeros(n, np.int32)
    L.ref_fastrandom_stream(seed, lo, hi, n, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def generate_coo_pairs(pseudo_nnz, max_size):
    """generate_coo_pairs_cpu (coo_utils.cpp:204-218): sorted unique pairs."""
    ptr = lib().orc_generate_coo_pairs(pseudo_nnz, max_size)
    if not ptr:
        raise ValueError("Incorrect generation")  # coo_utils.cpp:136-140
    return _take_pairs(ptr)


def generate_coo(pseudo_nnz, max_size):
    """generate_coo_cpu (coo_utils.cpp:220-239): draws in [0, max] inclusive."""
    return _take_pairs(lib().orc_generate_coo(pseudo_nnz, max_size))


def pairs_to_dcsr(rows, cols, nrows=0, ncols=0):
    """coo_pairs_to_dcsr_cpu (coo_utils.cpp:241-266)."""
    rows, cols = _u32(rows), _u32(cols)
    return _take_mat(lib().orc_pairs_to_dcsr(_p32(rows), _p32(cols), rows.size), nrows, ncols)


def spgemm_reference(a: Dcsr, b: Dcsr) -> Dcsr:
    """matrix_multiplication_cpu (coo_utils.cpp:470-514), linear find and all."""
    ar, aw, ac = _u32(a.rpt), _u32(a.rows), _u32(a.cols)
    br, bw, bc = _u32(b.rpt), _u32(b.rows), _u32(b.cols)
    ptr = lib().orc_spgemm_reference(_p32(ar), _p32(aw), _p32(ac), a.nzr, _p32(br), _p32(bw), _p32(bc), b.nzr)
    return _take_mat(ptr, a.nrows, b.ncols)


def spgemm(a: Dcsr, b: Dcsr, row_sel=None, threads=None) -> Dcsr:
    """Canonical-DCSR boolean product by Gustavson (same answer as above)."""
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    ar, aw, ac = _u32(a.rpt), _u32(a.rows), _u32(a.cols)
    br, bw, bc = _u32(b.rpt), _u32(b.rows), _u32(b.cols)
    if row_sel is not None:
        sel = _u32(row_sel)
        sp, ns = _p32(sel), sel.size
    else:
        sel, sp, ns = None, None, 0
    ptr = lib().orc_spgemm_gustavson(_p32(ar), _p32(aw), _p32(ac), a.nzr, _p32(br), _p32(bw), _p32(bc), b.nzr,
                                     b.nrows, b.ncols, sp, ns, threads)
    return _take_mat(ptr, a.nrows, b.ncols)


def spgemm_digest(a: Dcsr, b: Dcsr, r0=0, r1=None, threads=None):
    """Per-row digests (nnz, sum of splitmix64(col)) of rows [r0, r1) of
    A * B without materialising the product (checker for full-size runs)."""
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    r1 = a.nzr if r1 is None else r1
    cnt = np.zeros(max(0, r1 - r0), np.uint64)
    h = np.zeros(max(0, r1 - r0), np.uint64)
    ar, ac = _u32(a.rpt), _u32(a.cols)
    br, bw, bc = _u32(b.rpt), _u32(b.rows), _u32(b.cols)
    lib().orc_spgemm_digest(_p32(ar), _p32(ac), r0, r1, _p32(br), _p32(bw), _p32(bc), b.nzr, b.ncols, threads,
                            cnt.ctypes.data_as(u64p), h.ctypes.data_as(u64p))
    return cnt, h


def dcsr_row_digest(rpt, cols, base=0, r0=0, r1=None, threads=None):
    """Digests of stored rows [r0, r1) of a DCSR matrix from a chunk of its
    columns starting at column position `base`; returns (hashes, rows whose
    columns are not strictly ascending)."""
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    rpt, cols = _u32(rpt), _u32(cols)
    r1 = rpt.size - 1 if r1 is None else r1
    h = np.zeros(max(0, r1 - r0), np.uint64)
    bad = lib().orc_dcsr_row_digest(_p32(rpt), _p32(cols), base, r0, r1, threads, h.ctypes.data_as(u64p))
    return h, int(bad)


def dense_bool_mm(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    m, k = a.shape
    k2, n = b.shape
    assert k == k2
    c = np.zeros((m, n), np.uint8)
    lib().orc_dense_bool_mm(a.ctypes.data, b.ctypes.data, c.ctypes.data, m, k, n)
    return c


def workload(a: Dcsr, b: Dcsr):
    """get_workload (coo_utils.cpp:439-456): per-A-row product count."""
    out = np.zeros(a.nzr, np.uint64)
    ar, ac = _u32(a.rpt), _u32(a.cols)
    br, bw = _u32(b.rpt), _u32(b.rows)
    lib().orc_workload(_p32(ar), _p32(ac), a.nzr, _p32(br), _p32(bw), b.nzr, out.ctypes.data_as(u64p))
    return out


def coo_add(ar, ac, br, bc):
    """matrix_addition_cpu (coo_utils.cpp:316-324)."""
    ar, ac, br, bc = map(_u32, (ar, ac, br, bc))
    return _take_pairs(lib().orc_coo_add(_p32(ar), _p32(ac), ar.size, _p32(br), _p32(bc), br.size))


def kron(ar, ac, br, bc, b_nrows, b_ncols):
    """kronecker_product_cpu (coo_utils.cpp:327-342)."""
    ar, ac, br, bc = map(_u32, (ar, ac, br, bc))
    return _take_pairs(lib().orc_kron(_p32(ar), _p32(ac), ar.size, _p32(br), _p32(bc), br.size, b_nrows, b_ncols))


def exclusive_scan(a):
    """Exclusive scan plus total (test/src/procedures.cpp:24-31)."""
    a = _u32(a).copy()
    total = lib().orc_exclusive_scan_u32(_p32(a), a.size)
    return a, int(total)


def csr_from_pairs(rows, cols, nrows):
    """csr_cpu_from_pairs (matrices_conversions.cpp:158-177) -> (rpt, cols)."""
    rows, cols = _u32(rows), _u32(cols)
    m = _take_mat(lib().orc_csr_from_pairs(_p32(rows), _p32(cols), rows.size, nrows), with_rows=False)
    return m.rpt, m.cols


def transpose(m: Dcsr) -> Dcsr:
    """matrix_coo_cpu::transpose + coo_to_dcsr_cpu (cpu_matrices.hpp:89-103,
    coo_utils.cpp:268-294)."""
    rpt, rows, cols = _u32(m.rpt), _u32(m.rows), _u32(m.cols)
    return _take_mat(lib().orc_transpose(_p32(rpt), _p32(rows), _p32(cols), m.nzr), m.ncols, m.nrows)


def submatrix(m: Dcsr, i, j, nrows, ncols) -> Dcsr:
    """submatrix_cpu (cpu_matrix_operations.cpp:6-52)."""
    rpt, rows, cols = _u32(m.rpt), _u32(m.rows), _u32(m.cols)
    return _take_mat(lib().orc_submatrix(_p32(rpt), _p32(rows), _p32(cols), m.nzr, i, j, nrows, ncols),
                     nrows, ncols)


def reduce(m: Dcsr) -> Dcsr:
    """reduce (cpu_matrix_operations.cpp:63-76)."""
    rows = _u32(m.rows)
    return _take_mat(lib().orc_reduce(_p32(rows), m.nzr), m.nrows, 1)


def dcsr_to_csr_rpt(m: Dcsr, nrows):
    """dcsr_to_csr (matrices_conversions.cpp:247-271): the CSR row pointer."""
    out = np.zeros(nrows + 1, np.uint32)
    rpt, rows = _u32(m.rpt), _u32(m.rows)
    lib().orc_dcsr_to_csr_rpt(_p32(rpt), _p32(rows), m.nzr, nrows, _p32(out))
    return out


def csr_to_dcsr(csr_rpt, nrows):
    """csr_to_dcsr (matrices_conversions.cpp:211-244) -> (rpt, rows)."""
    r = _u32(csr_rpt)
    m = _take_mat(lib().orc_csr_to_dcsr(_p32(r), nrows))
    return m.rpt, m.rows


def coo_build(rows, cols, sorted_=True, no_duplicates=True):
    """matrix_coo host constructor (matrix_coo.cpp:17-47)."""
    rows, cols = _u32(rows), _u32(cols)
    return _take_pairs(lib().orc_coo_build(_p32(rows), _p32(cols), rows.size, int(sorted_), int(no_duplicates)))


# ---- MLP ----------------------------------------------------------------
def mlp_param_count(F, H1, H2):
    return int(lib().orc_mlp_param_count(_Dims(F, H1, H2)))


def mlp_forward(params, x, F, H1, H2):
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    y = np.zeros(x.shape[0], np.float32)
    lib().orc_mlp_forward(_Dims(F, H1, H2), p.ctypes.data, x.ctypes.data, x.shape[0], y.ctypes.data)
    return y


def mlp_loss_grad(params, x, t, F, H1, H2):
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    g = np.zeros(p.size, np.float64)
    loss = lib().orc_mlp_loss_grad(_Dims(F, H1, H2), p.ctypes.data, x.ctypes.data, t.ctypes.data,
                                   x.shape[0], g.ctypes.data)
    return float(loss), g


def adam_step(p, m, v, g, lr, b1, b2, eps, step):
    """In place on float32 arrays p, m, v with gradient g."""
    g = np.ascontiguousarray(g, np.float32)
    lib().orc_adam_step(p.size, p.ctypes.data, m.ctypes.data, v.ctypes.data, g.ctypes.data,
                        lr, b1, b2, eps, step)
