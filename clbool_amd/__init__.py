"""Python host mirror of clBool's API over the MI355X C-ABI (libclbool_amd.so).

Names and argument meaning follow the reference's include/clbool.hpp:
``create_controls``, ``matrix_dcsr`` / ``matrix_coo`` / ``matrix_csr``, the
conversions of src/common/matrices_conversions.hpp, ``dcsr.matrix_multiplication_hash``,
``dcsr.kronecker_product``, ``coo.matrix_addition``, ``csr.matrix_addition``,
``prefix_sum`` and ``fill_with``.  Operations return their output instead of
filling an out-parameter.  Errors raise ``ClboolError`` carrying clBool's
Status code (src/core/error.hpp:130-151).

Every call goes to hand-written HIP kernels; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
import types

import numpy as np

from . import _capi
from ._capi import Coo as _Coo, Csr as _Csr, Dcsr as _Dcsr, SpgemmOpts, SpgemmStats, Tuning

STATUS_NAMES = [
    "CLBOOL_NO_SUCH_PROGRAM", "CLBOOL_CREATE_KERNEL_ERROR", "CLBOOL_BUILD_PROGRAM_ERROR",
    "CLBOOL_INCOMPLETE_KERNEL", "CLBOOL_RUN_KERNEL_FAILURE", "CLBOOL_EVENT_WAITING_ERROR",
    "CLBOOL_CREATE_BUFFER_ERROR", "CLBOOL_COPY_BUFFER_ERROR", "CLBOOL_READ_BUFFER_ERROR",
    "CLBOOL_WRITE_BUFFER_ERROR", "CLBOOL_CREATE_CONTROLS_ERROR", "CLBOOL_INVALID_ARGUMENT",
    "CLBOOL_INVALID_VALUE", "CLBOOL_INITIALIZATION_ERROR",
]
SPGEMM_PLAIN = _capi.SPGEMM_PLAIN
SPGEMM_LEARNED = _capi.SPGEMM_LEARNED
SPGEMM_BOUND = _capi.SPGEMM_BOUND


class ClboolError(RuntimeError):
    """clbool::Exception: message plus Status."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.status_name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"[{self.status_name}] {message}")


def lib():
    return _capi.load()


def _check(rc: int):
    if rc != 0:
        msg = lib().clb_last_error()
        raise ClboolError(rc - 1, msg.decode() if msg else "")


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _p(a):
    return a.ctypes.data_as(_capi.u32p)


# ---------------------------------------------------------------------------
class Controls:
    """clbool::Controls (src/core/controls.hpp:12-28): one device, one stream."""

    def __init__(self, platform_id: int = 0, device_id: int = 0):
        h = C.c_void_p()
        _check(lib().clb_controls_create(platform_id, device_id, C.byref(h)))
        self.handle = h
        self.device_id = device_id

    def close(self):
        if getattr(self, "handle", None):
            lib().clb_controls_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().clb_controls_sync(self.handle))

    @property
    def max_wg_size(self):
        v = C.c_uint32()
        _check(lib().clb_controls_max_wg_size(self.handle, C.byref(v)))
        return v.value

    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().clb_get_tuning(self.handle, C.byref(t)))
        return t

    def set_tuning(self, t: Tuning):
        _check(lib().clb_set_tuning(self.handle, C.byref(t)))

    def last_stats(self) -> dict:
        s = SpgemmStats()
        _check(lib().clb_last_stats(self.handle, C.byref(s)))
        return s.as_dict()

    def pool_stats(self) -> dict:
        """Buffer-pool diagnostics: cached bytes, hipMalloc calls and bytes,
        cache drops after a failed hipMalloc, free and total device memory."""
        a = (C.c_uint64 * 6)()
        _check(lib().clb_pool_stats(self.handle, a))
        return {"cached_bytes": a[0], "mallocs": a[1], "malloc_bytes": a[2], "drops": a[3],
                "free_bytes": a[4], "total_bytes": a[5]}

    def trim(self):
        _check(lib().clb_pool_trim(self.handle))


def create_controls(platform_id: int = 0, device_id: int = 0) -> Controls:
    """create_controls (src/common/env.cpp:7-36)."""
    return Controls(platform_id, device_id)


def device_count() -> int:
    n = C.c_uint32()
    _check(lib().clb_device_count(C.byref(n)))
    return n.value


def show_devices() -> list:
    """show_devices (src/common/env.cpp:38-56), returned as dicts."""
    out = []
    for i in range(device_count()):
        info = _capi.DeviceInfo()
        _check(lib().clb_device_info_get(i, C.byref(info)))
        out.append({"id": i, "name": info.name.decode(), "arch": info.arch.decode(),
                    "compute_units": info.compute_units, "max_wg_size": info.max_wg_size,
                    "lds_bytes": info.lds_bytes, "global_mem_bytes": info.global_mem_bytes,
                    "mem_clock_khz": info.mem_clock_khz, "mem_bus_width": info.mem_bus_width})
    return out


class Buffer:
    """A reference-counted device buffer (replaces cl::Buffer)."""

    def __init__(self, handle, owned=True):
        self.handle = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        if not owned and self.handle:
            lib().clb_buffer_retain(self.handle)

    @classmethod
    def create(cls, controls: Controls, nbytes: int) -> "Buffer":
        h = C.c_void_p()
        _check(lib().clb_buffer_create(controls.handle, nbytes, C.byref(h)))
        return cls(h)

    @classmethod
    def from_array(cls, controls: Controls, arr) -> "Buffer":
        arr = np.ascontiguousarray(arr)
        h = C.c_void_p()
        _check(lib().clb_buffer_from_host(controls.handle, arr.ctypes.data, arr.nbytes, C.byref(h)))
        return cls(h)

    @property
    def nbytes(self):
        return int(lib().clb_buffer_size(self.handle)) if self.handle else 0

    @property
    def device_ptr(self):
        return lib().clb_buffer_device_ptr(self.handle)

    def read(self, controls: Controls, count: int, dtype=np.uint32, offset=0):
        out = np.empty(count, dtype=dtype)
        if count:
            _check(lib().clb_buffer_read(controls.handle, self.handle, offset, out.nbytes, out.ctypes.data))
        return out

    def write(self, controls: Controls, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        _check(lib().clb_buffer_write(controls.handle, self.handle, offset, arr.nbytes, arr.ctypes.data))

    def __del__(self):
        try:
            if self.handle:
                lib().clb_buffer_release(self.handle)
                self.handle = None
        except Exception:
            pass


def _take(h):
    return Buffer(h) if h else None


def _raw(b):
    return b.handle if b is not None else None


# ---------------------------------------------------------------------------
class matrix_dcsr:
    """clbool::matrix_dcsr (src/core/matrix_dcsr.hpp:11-84)."""

    def __init__(self, nrows=0, ncols=0, rpt=None, rows=None, cols=None, nnz=0, nzr=0):
        self.nrows, self.ncols, self.nnz, self.nzr = nrows, ncols, nnz, nzr
        self.rpt_gpu, self.rows_gpu, self.cols_gpu = rpt, rows, cols

    def empty(self):
        return self.nnz == 0

    def _c(self):
        return _Dcsr(_raw(self.rpt_gpu), _raw(self.rows_gpu), _raw(self.cols_gpu),
                     self.nrows, self.ncols, self.nnz, self.nzr)

    @classmethod
    def _from_c(cls, m: _Dcsr):
        return cls(m.nrows, m.ncols, _take(m.rpt), _take(m.rows), _take(m.cols), m.nnz, m.nzr)

    def to_host(self, controls):
        """matrix_dcsr_from_gpu (matrices_conversions.cpp:121-139): (rpt, rows, cols)."""
        if self.nnz == 0:
            return np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32)
        return (self.rpt_gpu.read(controls, self.nzr + 1), self.rows_gpu.read(controls, self.nzr),
                self.cols_gpu.read(controls, self.nnz))


class matrix_coo:
    """clbool::matrix_coo (src/core/matrix_coo.hpp:11-77)."""

    def __init__(self, nrows=0, ncols=0, rows=None, cols=None, nnz=0):
        self.nrows, self.ncols, self.nnz = nrows, ncols, nnz
        self.rows_gpu, self.cols_gpu = rows, cols

    def empty(self):
        return self.nnz == 0

    @classmethod
    def from_host(cls, controls, rows, cols, nrows, ncols, nnz=None, sorted=True, noDuplicates=True):
        """matrix_coo(controls, rows*, cols*, nrows, ncols, nnz, sorted, noDuplicates)."""
        rows, cols = _u32(rows), _u32(cols)
        nnz = rows.size if nnz is None else nnz
        m = _Coo()
        _check(lib().clb_coo_from_host(controls.handle, _p(rows), _p(cols), nrows, ncols, nnz,
                                       int(sorted), int(noDuplicates), C.byref(m)))
        return cls._from_c(m)

    @classmethod
    def from_buffers(cls, controls, nrows, ncols, nnz, rows: Buffer, cols: Buffer, sorted=True, noDuplicates=True):
        m = _Coo()
        _check(lib().clb_coo_from_buffers(controls.handle, nrows, ncols, nnz, _raw(rows), _raw(cols),
                                          int(sorted), int(noDuplicates), C.byref(m)))
        return cls._from_c(m)

    def _c(self):
        return _Coo(_raw(self.rows_gpu), _raw(self.cols_gpu), self.nrows, self.ncols, self.nnz)

    @classmethod
    def _from_c(cls, m: _Coo):
        return cls(m.nrows, m.ncols, _take(m.rows), _take(m.cols), m.nnz)

    def to_host(self, controls):
        """matrix_coo_from_gpu (matrices_conversions.cpp:142-156): (rows, cols)."""
        if self.nnz == 0:
            return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
        return self.rows_gpu.read(controls, self.nnz), self.cols_gpu.read(controls, self.nnz)


class matrix_csr:
    """clbool::matrix_csr (src/core/matrix_csr.hpp:7-44)."""

    def __init__(self, nrows=0, ncols=0, rpt=None, cols=None, nnz=0):
        self.nrows, self.ncols, self.nnz = nrows, ncols, nnz
        self.rpt_gpu, self.cols_gpu = rpt, cols

    def empty(self):
        return self.nnz == 0

    def _c(self):
        return _Csr(_raw(self.rpt_gpu), _raw(self.cols_gpu), self.nrows, self.ncols, self.nnz)

    @classmethod
    def _from_c(cls, m: _Csr):
        return cls(m.nrows, m.ncols, _take(m.rpt), _take(m.cols), m.nnz)

    def to_host(self, controls):
        """matrix_csr_from_gpu (matrices_conversions.cpp:273-285): (rpt, cols)."""
        if self.nnz == 0:
            return np.zeros(self.nrows + 1, np.uint32), np.zeros(0, np.uint32)
        return self.rpt_gpu.read(controls, self.nrows + 1), self.cols_gpu.read(controls, self.nnz)


# ---------------------------------------------------------------------------
def matrix_dcsr_from_cpu(controls, rpt, rows, cols, nrows, ncols=None) -> matrix_dcsr:
    """matrix_dcsr_from_cpu (matrices_conversions.cpp:93-108); ncols defaults to nrows."""
    rpt, rows, cols = _u32(rpt), _u32(rows), _u32(cols)
    ncols = nrows if ncols is None else ncols
    m = _Dcsr()
    _check(lib().clb_dcsr_from_host(controls.handle, _p(rpt), _p(rows), _p(cols), rows.size, cols.size,
                                    nrows, ncols, C.byref(m)))
    return matrix_dcsr._from_c(m)


def csr_from_cpu(controls, rpt, cols, nrows, ncols) -> matrix_csr:
    """csr_from_cpu (matrices_conversions.cpp:199-208)."""
    rpt, cols = _u32(rpt), _u32(cols)
    m = _Csr()
    _check(lib().clb_csr_from_host(controls.handle, _p(rpt), _p(cols), nrows, ncols, cols.size, C.byref(m)))
    return matrix_csr._from_c(m)


def coo_to_dcsr_shallow(controls, a: matrix_coo) -> matrix_dcsr:
    m, ac = _Dcsr(), a._c()
    _check(lib().clb_coo_to_dcsr_shallow(controls.handle, C.byref(ac), C.byref(m)))
    return matrix_dcsr._from_c(m)


def dcsr_to_coo_shallow(controls, a: matrix_dcsr) -> matrix_coo:
    m, ac = _Coo(), a._c()
    _check(lib().clb_dcsr_to_coo_shallow(controls.handle, C.byref(ac), C.byref(m)))
    return matrix_coo._from_c(m)


def dcsr_to_coo_deep(controls, a: matrix_dcsr) -> matrix_coo:
    m, ac = _Coo(), a._c()
    _check(lib().clb_dcsr_to_coo_deep(controls.handle, C.byref(ac), C.byref(m)))
    return matrix_coo._from_c(m)


def dcsr_to_csr(controls, a: matrix_dcsr) -> matrix_csr:
    m, ac = _Csr(), a._c()
    _check(lib().clb_dcsr_to_csr(controls.handle, C.byref(ac), C.byref(m)))
    return matrix_csr._from_c(m)


def csr_to_dcsr(controls, a: matrix_csr) -> matrix_dcsr:
    m, ac = _Dcsr(), a._c()
    _check(lib().clb_csr_to_dcsr(controls.handle, C.byref(ac), C.byref(m)))
    return matrix_dcsr._from_c(m)


def prefix_sum(controls, buf: Buffer, n: int) -> int:
    """prefix_sum (cl_operations.cpp:4-88): in-place exclusive scan, returns the total."""
    t = C.c_uint32()
    _check(lib().clb_prefix_sum(controls.handle, _raw(buf), n, C.byref(t)))
    return t.value


def fill_with(controls, buf: Buffer, n: int, value: int):
    """fill_with (cl_operations.cpp:90-100)."""
    _check(lib().clb_fill_with(controls.handle, _raw(buf), n, value))


def kernel_lookup(program: str, kernel: str):
    """KernelCache::get_kernel (kernel_cache.hpp:46-56) name check."""
    _check(lib().clb_kernel_lookup(program.encode(), kernel.encode()))


# ---------------------------------------------------------------------------
def _matrix_multiplication_hash(controls, a: matrix_dcsr, b: matrix_dcsr) -> matrix_dcsr:
    """dcsr::matrix_multiplication_hash (dcsr_matrix_multiplication_hash.cpp:51-119)."""
    m, ac, bc = _Dcsr(), a._c(), b._c()
    _check(lib().clb_dcsr_spgemm_hash(controls.handle, C.byref(ac), C.byref(bc), C.byref(m)))
    return matrix_dcsr._from_c(m)


def _matrix_multiplication(controls, a: matrix_dcsr, b: matrix_dcsr) -> matrix_dcsr:
    """dcsr::matrix_multiplication (dcsr_matrix_multiplication.cpp:26-168): the
    expand-sort-compress SpGEMM."""
    m, ac, bc = _Dcsr(), a._c(), b._c()
    _check(lib().clb_dcsr_spgemm_esc(controls.handle, C.byref(ac), C.byref(bc), C.byref(m)))
    return matrix_dcsr._from_c(m)


def _spgemm(controls, a: matrix_dcsr, b: matrix_dcsr, mode=SPGEMM_PLAIN, row_begin=0, row_end=0, timing=False):
    """The hash SpGEMM with a mode and an optional row block; returns (C, stats)."""
    m, ac, bc = _Dcsr(), a._c(), b._c()
    o = SpgemmOpts(mode, row_begin, row_end, int(timing))
    st = SpgemmStats()
    _check(lib().clb_dcsr_spgemm(controls.handle, C.byref(ac), C.byref(bc), C.byref(o), C.byref(m), C.byref(st)))
    return matrix_dcsr._from_c(m), st.as_dict()


class SpgemmPlan:
    """A * B taken in row blocks planned by the library (clb_spgemm_plan_*):
    the mode's sizing pass over all of A's stored rows, then consecutive
    blocks whose results hold at most block_bound nonzeros.  Iterating yields
    (C_block, stats) for every block; a learned block that exceeds the bound
    is cut again on the fly, so the block count can grow while iterating."""

    def __init__(self, controls, a: matrix_dcsr, b: matrix_dcsr, mode=SPGEMM_PLAIN, block_bound=0):
        self.controls = controls
        self.handle = C.c_void_p()
        ac, bc = a._c(), b._c()
        _check(lib().clb_spgemm_plan_create(controls.handle, C.byref(ac), C.byref(bc), mode, int(block_bound),
                                            C.byref(self.handle)))

    def info(self) -> dict:
        i = _capi.PlanInfo()
        _check(lib().clb_spgemm_plan_info_get(self.handle, C.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_ if f != "reserved"}

    @property
    def nblocks(self) -> int:
        return self.info()["nblocks"]

    def block(self, k: int):
        rb, re = C.c_uint32(), C.c_uint32()
        _check(lib().clb_spgemm_plan_block(self.handle, k, C.byref(rb), C.byref(re)))
        return rb.value, re.value

    def run(self, k: int, timing=False):
        m = _Dcsr()
        st = SpgemmStats()
        _check(lib().clb_spgemm_plan_run(self.controls.handle, self.handle, k, int(timing), C.byref(m),
                                         C.byref(st)))
        return matrix_dcsr._from_c(m), st.as_dict()

    def __iter__(self):
        k = 0
        while k < self.nblocks:
            yield self.run(k)
            k += 1

    def run_all(self, fn, timing=False):
        """Every block in order: fn(k, C_block, stats) as soon as block k is
        complete (a true return value stops the run).  In the one-pass modes
        the next block's passes overlap this block's last phase."""
        err = []

        def cb(user, k, blk, st):
            try:
                return 1 if fn(k, matrix_dcsr._from_c(blk.contents), st.contents.as_dict()) else 0
            except BaseException as e:  # reported after the call returns
                err.append(e)
                return 1
        c_fn = _capi.BLOCK_FN(cb)
        rc = lib().clb_spgemm_plan_run_all(self.controls.handle, self.handle, int(timing), c_fn, None)
        if err:
            raise err[0]
        _check(rc)

    def close(self):
        if self.handle:
            lib().clb_spgemm_plan_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _kronecker_product(controls, a: matrix_dcsr, b: matrix_dcsr) -> matrix_dcsr:
    """dcsr::kronecker_product (dcsr_kronecker_product.cpp:6-65)."""
    m, ac, bc = _Dcsr(), a._c(), b._c()
    _check(lib().clb_dcsr_kronecker(controls.handle, C.byref(ac), C.byref(bc), C.byref(m)))
    return matrix_dcsr._from_c(m)


def _transpose(controls, a: matrix_dcsr) -> matrix_dcsr:
    """dcsr::transpose (dcsr_transpose.cpp:5-17)."""
    m, ac = _Dcsr(), a._c()
    _check(lib().clb_dcsr_transpose(controls.handle, C.byref(ac), C.byref(m)))
    return matrix_dcsr._from_c(m)


def _submatrix(controls, a: matrix_dcsr, i: int, j: int, nrows: int, ncols: int) -> matrix_dcsr:
    """dcsr::submatrix (dcsr_submatrix.cpp:98-156)."""
    m, ac = _Dcsr(), a._c()
    _check(lib().clb_dcsr_submatrix(controls.handle, C.byref(ac), i, j, nrows, ncols, C.byref(m)))
    return matrix_dcsr._from_c(m)


def _reduce(controls, a: matrix_dcsr) -> matrix_dcsr:
    """dcsr::reduce (dcsr_reduce.cpp:5-31)."""
    m, ac = _Dcsr(), a._c()
    _check(lib().clb_dcsr_reduce(controls.handle, C.byref(ac), C.byref(m)))
    return matrix_dcsr._from_c(m)


def _digest(controls, a: matrix_dcsr) -> int:
    """64-bit device-side digest of a DCSR matrix (clb_dcsr_digest)."""
    h, ac = C.c_uint64(), a._c()
    _check(lib().clb_dcsr_digest(controls.handle, C.byref(ac), C.byref(h)))
    return h.value


def _dcsr_addition(controls, a: matrix_dcsr, b: matrix_dcsr) -> matrix_dcsr:
    m, ac, bc = _Dcsr(), a._c(), b._c()
    _check(lib().clb_dcsr_addition(controls.handle, C.byref(ac), C.byref(bc), C.byref(m)))
    return matrix_dcsr._from_c(m)


def _coo_matrix_addition(controls, a: matrix_coo, b: matrix_coo) -> matrix_coo:
    """coo::matrix_addition (coo_matrix_addition.cpp:32-73)."""
    m, ac, bc = _Coo(), a._c(), b._c()
    _check(lib().clb_coo_addition(controls.handle, C.byref(ac), C.byref(bc), C.byref(m)))
    return matrix_coo._from_c(m)


def _coo_sort_arrays(controls, rows: Buffer, cols: Buffer, n: int):
    """coo::sort_arrays (coo_initialization.cpp:6-32)."""
    _check(lib().clb_coo_sort_arrays(controls.handle, _raw(rows), _raw(cols), n))


def _csr_matrix_addition(controls, a: matrix_csr, b: matrix_csr) -> matrix_csr:
    """csr::matrix_addition (csr_addition.cpp:26-168)."""
    m, ac, bc = _Csr(), a._c(), b._c()
    _check(lib().clb_csr_addition(controls.handle, C.byref(ac), C.byref(bc), C.byref(m)))
    return matrix_csr._from_c(m)


dcsr = types.SimpleNamespace(matrix_multiplication_hash=_matrix_multiplication_hash,
                             matrix_multiplication=_matrix_multiplication,
                             kronecker_product=_kronecker_product, spgemm=_spgemm, plan=SpgemmPlan,
                             addition=_dcsr_addition, digest=_digest,
                             transpose=_transpose, submatrix=_submatrix, reduce=_reduce)
coo = types.SimpleNamespace(matrix_addition=_coo_matrix_addition, sort_arrays=_coo_sort_arrays)
csr = types.SimpleNamespace(matrix_addition=_csr_matrix_addition)


# ---------------------------------------------------------------------------
# Learned row-size predictor (no counterpart in the reference; see
# include/clbool_c.h).  The kernels run on the GPU; these wrappers only move
# arguments across the C-ABI.
MLP_PARAMS = _capi.MLP_PARAMS


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def predictor_reset(controls, seed=1):
    _check(lib().clb_predictor_reset(controls.handle, seed))


def predictor_harvest(controls, a: matrix_dcsr, b: matrix_dcsr, max_rows=1 << 20) -> int:
    """Append (features, exact row size) samples of A * B; returns the total."""
    n = C.c_uint64()
    ac, bc = a._c(), b._c()
    _check(lib().clb_predictor_harvest(controls.handle, C.byref(ac), C.byref(bc), max_rows, C.byref(n)))
    return n.value


def predictor_train(controls, epochs=20, batch=4096, lr=3e-3, seed=1) -> dict:
    o = _capi.TrainOpts(epochs, batch, lr, seed)
    s = _capi.TrainStats()
    _check(lib().clb_predictor_train(controls.handle, C.byref(o), C.byref(s)))
    return {f: getattr(s, f) for f, _ in s._fields_}


def predictor_get_params(controls) -> np.ndarray:
    p = np.zeros(MLP_PARAMS, np.float32)
    _check(lib().clb_predictor_get_params(controls.handle, p.ctypes.data, MLP_PARAMS))
    return p


def predictor_set_params(controls, params):
    p = _f32(params)
    _check(lib().clb_predictor_set_params(controls.handle, p.ctypes.data, p.size))


def predictor_predict(controls, a: matrix_dcsr, b: matrix_dcsr) -> np.ndarray:
    out = np.zeros(a.nzr, np.float32)
    ac, bc = a._c(), b._c()
    _check(lib().clb_predictor_predict(controls.handle, C.byref(ac), C.byref(bc), out.ctypes.data))
    return out


def mlp_forward(controls, params, x) -> np.ndarray:
    p, xx = _f32(params), _f32(x)
    y = np.zeros(xx.shape[0], np.float32)
    _check(lib().clb_mlp_forward(controls.handle, p.ctypes.data, xx.ctypes.data, xx.shape[0], y.ctypes.data))
    return y


def mlp_loss_grad(controls, params, x, t):
    p, xx, tt = _f32(params), _f32(x), _f32(t)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = C.c_float()
    _check(lib().clb_mlp_loss_grad(controls.handle, p.ctypes.data, xx.ctypes.data, tt.ctypes.data, xx.shape[0],
                                   g.ctypes.data, C.byref(loss)))
    return float(loss.value), g


def mlp_adam(controls, params, m, v, g, lr, b1, b2, eps, step):
    """One Adam step in place on float32 arrays params, m, v."""
    gg = _f32(g)
    _check(lib().clb_mlp_adam(controls.handle, params.ctypes.data, m.ctypes.data, v.ctypes.data, gg.ctypes.data,
                              lr, b1, b2, eps, step))


# ---------------------------------------------------------------------------
# Closure step of CFPQ/RPQ evaluation (see include/clbool_c.h).
CLOSURE_PER_TRIPLE = _capi.CLOSURE_PER_TRIPLE
CLOSURE_BATCHED = _capi.CLOSURE_BATCHED


def closure_step(controls, mats: list, triples, mode=CLOSURE_BATCHED, spgemm_mode=SPGEMM_PLAIN,
                 block_bound=0) -> dict:
    """X |= Y * Z for every (X, Y, Z) index triple; replaces grown matrices in `mats`."""
    arr = (_Dcsr * len(mats))()
    for k, m in enumerate(mats):
        arr[k] = m._c()
        for b in (m.rpt_gpu, m.rows_gpu, m.cols_gpu):  # the struct owns one reference
            if b is not None:
                lib().clb_buffer_retain(b.handle)
    flat = (C.c_uint32 * (3 * len(triples)))(*[int(v) for t in triples for v in t])
    o = _capi.ClosureOpts(mode, spgemm_mode, block_bound)
    st = _capi.ClosureStats()
    rc = lib().clb_closure_step(controls.handle, arr, len(mats), flat, len(triples), C.byref(o), C.byref(st))
    for k in range(len(mats)):  # adopt what the struct holds now (old or new matrix)
        mats[k] = matrix_dcsr._from_c(arr[k])
    _check(rc)
    return {f: getattr(st, f) for f, _ in st._fields_}
