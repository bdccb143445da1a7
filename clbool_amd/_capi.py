"""ctypes declarations of include/clbool_c.h (libclbool_amd.so).

The library is loaded from this package directory only; when it is missing
the import fails loudly (there is no CPU fallback of any operation).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclbool_amd.so")

u32p = C.POINTER(C.c_uint32)
buf_t = C.c_void_p
ctl_t = C.c_void_p

CLB_MAX_BINS = 32
SPGEMM_PLAIN = 0
SPGEMM_LEARNED = 1
SPGEMM_BOUND = 2
TUNE_BITMAP_PRED_CAPS = 0  # index into Tuning.reserved


class Dcsr(C.Structure):
    _fields_ = [("rpt", buf_t), ("rows", buf_t), ("cols", buf_t),
                ("nrows", C.c_uint32), ("ncols", C.c_uint32), ("nnz", C.c_uint32), ("nzr", C.c_uint32)]


class Coo(C.Structure):
    _fields_ = [("rows", buf_t), ("cols", buf_t),
                ("nrows", C.c_uint32), ("ncols", C.c_uint32), ("nnz", C.c_uint32)]


class Csr(C.Structure):
    _fields_ = [("rpt", buf_t), ("cols", buf_t),
                ("nrows", C.c_uint32), ("ncols", C.c_uint32), ("nnz", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("hash_max_slots", C.c_uint32), ("bitmap_min_products", C.c_uint32),
                ("bitmap_slot_ratio", C.c_uint32), ("tiny_max_products", C.c_uint32),
                ("learned_margin_pct", C.c_uint32), ("emit_lane_bits_min", C.c_uint32),
                ("dump_bits_per_chunk", C.c_uint32), ("bitmap_persistent", C.c_uint32),
                ("reserved", C.c_uint32 * 8)]


class SpgemmOpts(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("row_begin", C.c_uint32), ("row_end", C.c_uint32),
                ("timing", C.c_uint32)]


class SpgemmStats(C.Structure):
    _fields_ = [("products", C.c_uint64), ("nnz", C.c_uint64), ("nzr", C.c_uint32), ("rows", C.c_uint32),
                ("fallback_rows", C.c_uint32), ("n_bins", C.c_uint32),
                ("bin_rows", C.c_uint32 * CLB_MAX_BINS), ("bin_products", C.c_uint64 * CLB_MAX_BINS),
                ("bin_a_nnz", C.c_uint64 * CLB_MAX_BINS), ("bin_out_nnz", C.c_uint64 * CLB_MAX_BINS),
                ("pred_abs_log2_err", C.c_double),
                ("ms_setup", C.c_double), ("ms_symbolic", C.c_double), ("ms_numeric", C.c_double),
                ("ms_assemble", C.c_double), ("ms_total", C.c_double),
                ("ms_kernel", C.c_double * 16), ("launches", C.c_uint32 * 16), ("capacity", C.c_uint64),
                ("dense_rows", C.c_uint64), ("dense_nnz", C.c_uint64)]

    _ARRAYS = ("bin_rows", "bin_products", "bin_a_nnz", "bin_out_nnz", "ms_kernel", "launches")

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in self._ARRAYS}
        for f in self._ARRAYS:
            d[f] = list(getattr(self, f))
        return d


KERNEL_PASSES = ["sym", "num", "onepass", "fallback"]
KERNEL_KINDS = ["tiny", "hash", "bitmap", "emit"]
KERNEL_FAMILIES = [f"{p}_{k}" for p in KERNEL_PASSES for k in KERNEL_KINDS]
# the re-run has no emission launch; its slot holds the one pass's compaction
KERNEL_FAMILIES[15] = "onepass_compact"


BLOCK_FN = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_uint32, C.POINTER(Dcsr), C.POINTER(SpgemmStats))


class PlanInfo(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("nblocks", C.c_uint32), ("replans", C.c_uint32), ("reserved", C.c_uint32),
                ("products", C.c_uint64), ("planned_nnz", C.c_uint64), ("ms_plan", C.c_double)]


class ClosureOpts(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("spgemm_mode", C.c_uint32), ("block_bound", C.c_uint32),
                ("reserved", C.c_uint32)]


class ClosureStats(C.Structure):
    _fields_ = [("products", C.c_uint64), ("added", C.c_uint64), ("changed", C.c_uint32),
                ("spgemm_calls", C.c_uint32), ("ms", C.c_double), ("max_block_bound", C.c_uint64),
                ("ms_products", C.c_double), ("ms_unions", C.c_double), ("fallback_rows", C.c_uint64),
                ("dense_rows", C.c_uint64)]


CLOSURE_PER_TRIPLE = 0
CLOSURE_BATCHED = 1
MLP_PARAMS = 1281


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("seed", C.c_uint32)]


class TrainStats(C.Structure):
    _fields_ = [("loss_first", C.c_double), ("loss_last", C.c_double), ("samples", C.c_uint64),
                ("steps", C.c_uint32), ("ms", C.c_double)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("compute_units", C.c_uint32),
                ("max_wg_size", C.c_uint32), ("lds_bytes", C.c_uint32), ("global_mem_bytes", C.c_uint64),
                ("clock_khz", C.c_uint32), ("mem_clock_khz", C.c_uint32), ("mem_bus_width", C.c_uint32)]


# (name, restype, argtypes) for every symbol of include/clbool_c.h
SIGNATURES = [
    ("clb_last_error", C.c_char_p, []),
    ("clb_device_count", C.c_int32, [u32p]),
    ("clb_controls_create", C.c_int32, [C.c_uint32, C.c_uint32, C.POINTER(ctl_t)]),
    ("clb_controls_destroy", C.c_int32, [ctl_t]),
    ("clb_device_info_get", C.c_int32, [C.c_uint32, C.POINTER(DeviceInfo)]),
    ("clb_controls_sync", C.c_int32, [ctl_t]),
    ("clb_controls_max_wg_size", C.c_int32, [ctl_t, u32p]),
    ("clb_get_tuning", C.c_int32, [ctl_t, C.POINTER(Tuning)]),
    ("clb_set_tuning", C.c_int32, [ctl_t, C.POINTER(Tuning)]),
    ("clb_last_stats", C.c_int32, [ctl_t, C.POINTER(SpgemmStats)]),
    ("clb_pool_trim", C.c_int32, [ctl_t]),
    ("clb_pool_stats", C.c_int32, [ctl_t, C.POINTER(C.c_uint64)]),
    ("clb_buffer_create", C.c_int32, [ctl_t, C.c_uint64, C.POINTER(buf_t)]),
    ("clb_buffer_from_host", C.c_int32, [ctl_t, C.c_void_p, C.c_uint64, C.POINTER(buf_t)]),
    ("clb_buffer_read", C.c_int32, [ctl_t, buf_t, C.c_uint64, C.c_uint64, C.c_void_p]),
    ("clb_buffer_write", C.c_int32, [ctl_t, buf_t, C.c_uint64, C.c_uint64, C.c_void_p]),
    ("clb_buffer_copy", C.c_int32, [ctl_t, buf_t, C.c_uint64, buf_t, C.c_uint64, C.c_uint64]),
    ("clb_buffer_retain", C.c_int32, [buf_t]),
    ("clb_buffer_release", C.c_int32, [buf_t]),
    ("clb_buffer_device_ptr", C.c_void_p, [buf_t]),
    ("clb_buffer_size", C.c_uint64, [buf_t]),
    ("clb_dcsr_release", C.c_int32, [C.POINTER(Dcsr)]),
    ("clb_coo_release", C.c_int32, [C.POINTER(Coo)]),
    ("clb_csr_release", C.c_int32, [C.POINTER(Csr)]),
    ("clb_coo_from_host", C.c_int32, [ctl_t, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                      C.POINTER(Coo)]),
    ("clb_coo_from_buffers", C.c_int32, [ctl_t, C.c_uint32, C.c_uint32, C.c_uint32, buf_t, buf_t, C.c_int,
                                         C.c_int, C.POINTER(Coo)]),
    ("clb_coo_sort_arrays", C.c_int32, [ctl_t, buf_t, buf_t, C.c_uint32]),
    ("clb_dcsr_from_host", C.c_int32, [ctl_t, u32p, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(Dcsr)]),
    ("clb_csr_from_host", C.c_int32, [ctl_t, u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Csr)]),
    ("clb_coo_to_dcsr_shallow", C.c_int32, [ctl_t, C.POINTER(Coo), C.POINTER(Dcsr)]),
    ("clb_dcsr_to_coo_shallow", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Coo)]),
    ("clb_dcsr_to_coo_deep", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Coo)]),
    ("clb_dcsr_to_csr", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Csr)]),
    ("clb_csr_to_dcsr", C.c_int32, [ctl_t, C.POINTER(Csr), C.POINTER(Dcsr)]),
    ("clb_dcsr_spgemm_hash", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.POINTER(Dcsr)]),
    ("clb_dcsr_spgemm", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.POINTER(SpgemmOpts),
                                    C.POINTER(Dcsr), C.POINTER(SpgemmStats)]),
    ("clb_dcsr_spgemm_esc", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.POINTER(Dcsr)]),
    ("clb_spgemm_plan_create", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.c_uint32, C.c_uint64,
                                           C.POINTER(C.c_void_p)]),
    ("clb_spgemm_plan_info_get", C.c_int32, [C.c_void_p, C.POINTER(PlanInfo)]),
    ("clb_spgemm_plan_block", C.c_int32, [C.c_void_p, C.c_uint32, u32p, u32p]),
    ("clb_spgemm_plan_run", C.c_int32, [ctl_t, C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(Dcsr),
                                        C.POINTER(SpgemmStats)]),
    ("clb_spgemm_plan_run_all", C.c_int32, [ctl_t, C.c_void_p, C.c_uint32, BLOCK_FN, C.c_void_p]),
    ("clb_spgemm_plan_release", C.c_int32, [C.c_void_p]),
    ("clb_dcsr_kronecker", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.POINTER(Dcsr)]),
    ("clb_dcsr_transpose", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr)]),
    ("clb_dcsr_submatrix", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(Dcsr)]),
    ("clb_dcsr_reduce", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr)]),
    ("clb_dcsr_digest", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(C.c_uint64)]),
    ("clb_coo_addition", C.c_int32, [ctl_t, C.POINTER(Coo), C.POINTER(Coo), C.POINTER(Coo)]),
    ("clb_csr_addition", C.c_int32, [ctl_t, C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Csr)]),
    ("clb_dcsr_addition", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.POINTER(Dcsr)]),
    ("clb_closure_step", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32,
                                     C.POINTER(ClosureOpts), C.POINTER(ClosureStats)]),
    ("clb_predictor_reset", C.c_int32, [ctl_t, C.c_uint32]),
    ("clb_predictor_harvest", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.c_uint32,
                                          C.POINTER(C.c_uint64)]),
    ("clb_predictor_train", C.c_int32, [ctl_t, C.POINTER(TrainOpts), C.POINTER(TrainStats)]),
    ("clb_predictor_get_params", C.c_int32, [ctl_t, C.c_void_p, C.c_uint32]),
    ("clb_predictor_set_params", C.c_int32, [ctl_t, C.c_void_p, C.c_uint32]),
    ("clb_predictor_predict", C.c_int32, [ctl_t, C.POINTER(Dcsr), C.POINTER(Dcsr), C.c_void_p]),
    ("clb_mlp_forward", C.c_int32, [ctl_t, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("clb_mlp_loss_grad", C.c_int32, [ctl_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                      C.c_void_p]),
    ("clb_mlp_adam", C.c_int32, [ctl_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                                 C.c_float, C.c_float, C.c_uint32]),
    ("clb_prefix_sum", C.c_int32, [ctl_t, buf_t, C.c_uint32, u32p]),
    ("clb_prefix_sum_noalloc", C.c_int32, [ctl_t, buf_t, C.c_uint32, u32p]),
    ("clb_fill_with", C.c_int32, [ctl_t, buf_t, C.c_uint32, C.c_uint32]),
    ("clb_kernel_lookup", C.c_int32, [C.c_char_p, C.c_char_p]),
]

_lib = None


def load():
    """Load libclbool_amd.so with every declared symbol bound."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python clbool_amd/build.py` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols(header_path=None):
    """Function names declared in include/*.h (for the export check)."""
    import glob
    import re
    root = os.path.dirname(HERE)
    paths = [header_path] if header_path else sorted(glob.glob(os.path.join(root, "include", "*.h")))
    names = []
    for p in paths:
        txt = open(p).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?\*?\s*\b(clb_[a-z0-9_]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))
