"""Python binding (ctypes) of libxplus_amd.so, the MI355X Jacobi-PCG library.

The library is C++/HIP; this module only marshals numpy arrays across the
C-ABI declared in include/xplus_capi.h.  Importing it loads the built
library and raises ImportError if it is missing -- there is no Python or CPU
fallback for any computation.

Layout of the API (mirrors the C-ABI):
  * drop-in kernel surface: spmv_csr_kernel, spmv_blocked_kernel,
    init_pcg_kernel, dot_kernel, update_xrz_kernel, update_p_kernel
    (archived/include/cg_kernels.hpp:9-54), on numpy float64 arrays;
  * Csr: host CSR + b + x0 (loader CsrDataset::load, generators);
  * Matrix: device-resident SELL-C-sigma matrix;
  * Pcg: graph-captured Jacobi-PCG on one system or a batch.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "lib" / "libxplus_amd.so"
CLI_PATH = PKG_DIR / "lib" / "xplus_local"
CUPER_CLI_PATH = PKG_DIR / "lib" / "xplus_cuper"
DEVICE_CLI_PATH = PKG_DIR / "lib" / "xplus_device"

if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C project-xplus_amd)")

_lib = C.CDLL(str(LIB_PATH))

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_vp = C.c_void_p

CONVERGED, MAX_ITER, BREAKDOWN, STAGNATION, RUNNING = 0, 1, 2, 3, -1
STATUS_NAMES = {0: "converged", 1: "max_iter", 2: "breakdown", 3: "stagnation", -1: "running"}


class MatrixOpts(C.Structure):
    _fields_ = [("sigma", C.c_int), ("keep_lo", C.c_int), ("keep64", C.c_int), ("grid", C.c_int),
                ("spmv_variant", C.c_int), ("block_rows", C.c_int), ("blocking", C.c_int),
                ("device_build", C.c_int)]


class MatrixInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("n_pad", C.c_int), ("nslices", C.c_int), ("nsys", C.c_int),
                ("sigma", C.c_int), ("idx16", C.c_int), ("grid", C.c_int),
                ("nnz", C.c_longlong), ("stored", C.c_longlong), ("bytes_spmv", C.c_longlong)]


class SolveOpts(C.Structure):
    _fields_ = [("tau", C.c_double), ("rel_tol", C.c_double), ("max_iters", C.c_int),
                ("graph_iters", C.c_int), ("use_graph", C.c_int), ("trace_cap", C.c_int),
                ("stall_limit", C.c_int), ("fused", C.c_int), ("accurate", C.c_int), ("full_work_map", C.c_int)]


class TuneReport(C.Structure):
    _fields_ = [("sigma", C.c_int), ("spmv_variant", C.c_int), ("grid", C.c_int),
                ("us_per_iter", C.c_double), ("us_per_iter_default", C.c_double), ("candidates", C.c_int)]


class Status(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("rz", C.c_double), ("rr", C.c_double),
                ("pap", C.c_double), ("alpha", C.c_double), ("beta", C.c_double), ("rr0", C.c_double),
                ("tau", C.c_double), ("fallback", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def _sig(name, restype, argtypes):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


# drop-in surface (void returns, silent no-op for n < 0)
_spmv_csr = _sig("spmv_csr_kernel", None, [_i32p, _i32p, _f64p, _f64p, _f64p, C.c_int])
_spmv_blocked = _sig("spmv_blocked_kernel", None, [_i32p, _i32p, _f64p, _f64p, _f64p, C.c_int])
_init_pcg = _sig("init_pcg_kernel", None, [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int])
_dot = _sig("dot_kernel", None, [_f64p, _f64p, _f64p, C.c_int])
_update_xrz = _sig("update_xrz_kernel", None,
                   [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int])
_update_p = _sig("update_p_kernel", None, [_f64p, _f64p, C.c_double, C.c_int])
_legacy_reset = _sig("xp_legacy_reset", None, [])
_last_error = _sig("xp_last_error", C.c_int, [C.c_char_p, C.c_int])
device_count = _sig("xp_device_count", C.c_int, [])

_m_create = _sig("xp_matrix_create", C.c_int,
                 [C.c_int, _i32p, _i32p, _f64p, C.POINTER(MatrixOpts), C.POINTER(_vp)])
_m_create_batch = _sig("xp_matrix_create_batch", C.c_int,
                       [C.c_int, _i32p, _i32p, _i32p, _f64p, C.POINTER(MatrixOpts), C.POINTER(_vp)])
_m_destroy = _sig("xp_matrix_destroy", None, [_vp])
_m_info = _sig("xp_matrix_get_info", C.c_int, [_vp, C.POINTER(MatrixInfo)])
_m_perm = _sig("xp_matrix_permutation", C.c_int, [_vp, _i32p])
_m_set_grid = _sig("xp_matrix_set_grid", C.c_int, [_vp, C.c_int])
_m_set_variant = _sig("xp_matrix_set_variant", C.c_int, [_vp, C.c_int])
_m_autotune = _sig("xp_matrix_autotune", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(TuneReport)])
_spmv_f32 = _sig("xp_spmv_f32", C.c_int, [_vp, _vp, _vp, _vp])
_spmv_f64 = _sig("xp_spmv_f64", C.c_int, [_vp, _vp, _vp, _vp])
_to_solve = _sig("xp_to_solve_order", C.c_int, [_vp, _vp, _vp, _vp])
_from_solve = _sig("xp_from_solve_order", C.c_int, [_vp, _vp, _vp, _vp])
_csr_spmv_f64 = _sig("xp_csr_spmv_f64", C.c_int, [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p])
_dev_alloc = _sig("xp_device_alloc", C.c_int, [C.c_ulonglong, C.POINTER(_vp)])
_dev_free = _sig("xp_device_free", C.c_int, [_vp])
_htod = _sig("xp_memcpy_htod", C.c_int, [_vp, _vp, C.c_ulonglong])
_dtoh = _sig("xp_memcpy_dtoh", C.c_int, [_vp, _vp, C.c_ulonglong])
_dtod = _sig("xp_memcpy_dtod", C.c_int, [_vp, _vp, C.c_ulonglong])
_dev_sync = _sig("xp_device_synchronize", C.c_int, [])
_pool_trim = _sig("xp_device_pool_trim", C.c_int, [])
_pool_bytes = _sig("xp_device_pool_bytes", C.c_int, [C.POINTER(C.c_ulonglong)])
_op_init = _sig("xp_init_pcg_f32", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp])
_op_dot = _sig("xp_dot_f32", C.c_int, [_vp, _vp, _vp, _vp, _vp])
_op_xrz = _sig("xp_update_xrz_f32", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp])
_op_p = _sig("xp_update_p_f32", C.c_int, [_vp, _vp, _vp, C.c_double, _vp])
_spmv_time = _sig("xp_spmv_f32_time", C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)])
_stream_time = _sig("xp_stream_time", C.c_int, [C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)])
_spmv_time_cold = _sig("xp_spmv_f32_time_cold", C.c_int, [_vp, C.c_int, C.c_ulonglong, C.POINTER(C.c_double)])

_p_create = _sig("xp_pcg_create", C.c_int, [_vp, C.POINTER(SolveOpts), C.POINTER(_vp)])
_p_destroy = _sig("xp_pcg_destroy", None, [_vp])
_p_set_rhs = _sig("xp_pcg_set_rhs", C.c_int, [_vp, _f64p, _vp])
_p_run = _sig("xp_pcg_run", C.c_int, [_vp])
_p_run_fixed = _sig("xp_pcg_run_fixed", C.c_int, [_vp, C.c_int])
_p_run_refined = _sig("xp_pcg_run_refined", C.c_int,
                      [_vp, C.c_double, C.c_double, C.c_int, _f64p, C.POINTER(C.c_int), C.POINTER(C.c_int)])
_p_run_reliable = _sig("xp_pcg_run_reliable", C.c_int,
                       [_vp, C.c_double, C.c_double, _f64p, C.POINTER(C.c_int), C.POINTER(C.c_int)])
_p_kernel_times = _sig("xp_pcg_kernel_times", C.c_int, [_vp, C.c_int, _f64p, C.POINTER(C.c_int)])
_p_kernel_times_split = _sig("xp_pcg_kernel_times_split", C.c_int, [_vp, C.c_int, _f64p, C.POINTER(C.c_int)])
_p_status = _sig("xp_pcg_get_status", C.c_int, [_vp, C.c_int, C.POINTER(Status)])
_p_get_x = _sig("xp_pcg_get_x", C.c_int, [_vp, _f64p])
_p_trace = _sig("xp_pcg_get_trace", C.c_int, [_vp, C.c_int, _f64p, C.c_int])
_p_true_res = _sig("xp_pcg_true_residual", C.c_int, [_vp, _f64p])
_p_timing = _sig("xp_pcg_last_timing", C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int)])

_p_set_precond = _sig("xp_pcg_set_preconditioner", C.c_int, [_vp, C.c_int, _vp, C.c_longlong])
_p_block_op = _sig("xp_pcg_block_operator", C.c_int, [_vp, _f64p])


class JacobiOpts(C.Structure):
    _fields_ = [("tau", C.c_double), ("max_iters", C.c_int), ("use_tau", C.c_int), ("graph_iters", C.c_int),
                ("use_graph", C.c_int)]


class JacobiStatus(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("diff", C.c_double)]


_j_create = _sig("xp_jacobi_create", C.c_int, [_vp, C.POINTER(JacobiOpts), C.POINTER(_vp)])
_j_destroy = _sig("xp_jacobi_destroy", None, [_vp])
_j_set_rhs = _sig("xp_jacobi_set_rhs", C.c_int, [_vp, _f64p, _vp])
_j_run = _sig("xp_jacobi_run", C.c_int, [_vp])
_j_status = _sig("xp_jacobi_get_status", C.c_int, [_vp, C.c_int, C.POINTER(JacobiStatus)])
_j_get_x = _sig("xp_jacobi_get_x", C.c_int, [_vp, _f64p])
_j_timing = _sig("xp_jacobi_last_timing", C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int)])
_j_kernel_time = _sig("xp_jacobi_kernel_time", C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)])
_rhs_ones = _sig("xp_rhs_ones_f32", C.c_int, [C.c_int, _i32p, _i32p, _f64p, _f64p])
_p_set_mlp = _sig("xp_pcg_set_mlp_preconditioner", C.c_int, [_vp, _vp])
class MultilevelOpts(C.Structure):
    _fields_ = [("coarse_max", C.c_int), ("omega", C.c_double), ("gamma", C.c_double),
                ("coarse_blocking", C.c_int), ("consecutive_rows", C.c_int), ("theta", C.c_double)]


_p_set_ml = _sig("xp_pcg_set_multilevel", C.c_int, [_vp, C.POINTER(MultilevelOpts)])
_p_ml_info = _sig("xp_pcg_multilevel_info", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), _i32p,
                                                     np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                                     C.POINTER(C.c_double)])
_p_ml_level = _sig("xp_pcg_multilevel_level", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp])
_p_ml_slots = _sig("xp_pcg_multilevel_slots", C.c_int, [_vp, C.c_int, _vp, C.POINTER(C.c_int)])
_p_ml_rounds = _sig("xp_pcg_multilevel_rounds", C.c_int, [_vp, C.c_int, _i32p, C.POINTER(C.c_int)])
_p_ml_phases = _sig("xp_pcg_multilevel_phases", C.c_int, [_vp, C.c_int, _f64p, C.POINTER(C.c_int)])
_p_ml_apply = _sig("xp_pcg_multilevel_apply", C.c_int, [_vp, _vp, _vp])
_p_ml_time = _sig("xp_pcg_multilevel_time", C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)])
_m_num_blocks = _sig("xp_matrix_num_blocks", C.c_int, [_vp])
_m_block_features = _sig("xp_matrix_block_features", C.c_int, [_vp, _f32p, _f32p])
_mlp_create = _sig("xp_mlp_create", C.c_int, [C.c_ulonglong, C.POINTER(_vp)])
_mlp_destroy = _sig("xp_mlp_destroy", None, [_vp])
_mlp_nparams = _sig("xp_mlp_num_params", C.c_int, [])
_mlp_get = _sig("xp_mlp_get_params", C.c_int, [_vp, _f32p])
_mlp_set = _sig("xp_mlp_set_params", C.c_int, [_vp, _f32p])
_bs_from = _sig("xp_blockset_from_csr", C.c_int, [C.c_int, _i32p, _i32p, _f64p, C.POINTER(_vp)])
_bs_from_slots = _sig("xp_blockset_from_csr_slots", C.c_int,
                      [C.c_int, _i32p, _i32p, _f64p, _i32p, C.c_int, C.c_double, C.POINTER(_vp)])
_bs_from_blocked = _sig("xp_blockset_from_csr_blocked", C.c_int,
                        [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_double, C.POINTER(_vp)])
_strength_blocking = _sig("xp_strength_blocking", C.c_int,
                          [C.c_int, _i32p, _i32p, _f64p, _i32p, C.POINTER(C.c_int)])
_aggregate_blocking = _sig("xp_aggregate_blocking", C.c_int,
                           [C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, C.POINTER(C.c_int)])
_m_block_rows = _sig("xp_matrix_block_rows", C.c_int, [_vp, _i32p])
_bs_append = _sig("xp_blockset_append", C.c_int, [_vp, _vp])
_bs_size = _sig("xp_blockset_size", C.c_int, [_vp])
_bs_copy = _sig("xp_blockset_copy", C.c_int, [_vp, _f32p, _f32p, _f32p, _f32p])
_bs_free = _sig("xp_blockset_free", None, [_vp])
_mlp_loss_grad = _sig("xp_mlp_loss_grad", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_double), _f32p])
_mlp_forward = _sig("xp_mlp_forward", C.c_int, [_vp, _vp, C.c_int, C.c_int, _f32p])
_mlp_blocks = _sig("xp_mlp_blocks", C.c_int, [_vp, _vp, _f32p])
_mlp_train = _sig("xp_mlp_train", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _f64p])

_ds_load = _sig("xp_dataset_load", C.c_int, [C.c_char_p, C.POINTER(_vp)])
_ds_save = _sig("xp_dataset_save", C.c_int, [_vp, C.c_char_p])
_gen2d = _sig("xp_gen_diffusion2d", C.c_int, [C.c_int, C.c_int, C.c_double, C.c_ulonglong, C.POINTER(_vp)])
_gen3d = _sig("xp_gen_diffusion3d", C.c_int, [C.c_int, C.c_double, C.c_ulonglong, C.POINTER(_vp)])
_genfem = _sig("xp_gen_fem2d", C.c_int,
               [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_ulonglong, C.POINTER(_vp)])
_csr_from = _sig("xp_csr_from_arrays", C.c_int, [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.POINTER(_vp)])
_csr_info = _sig("xp_csr_info", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)])
_csr_copy = _sig("xp_csr_copy", C.c_int, [_vp, _i32p, _i32p, _f64p, _f64p, _f64p])
_csr_free = _sig("xp_csr_free", None, [_vp])

_u64pp = C.POINTER(C.POINTER(C.c_ulonglong))
_cu_pack = _sig("xp_cuper_pack", C.c_int, [C.c_int, _i32p, _i32p, _f64p, C.POINTER(_vp)])
_cu_info = _sig("xp_cuper_packed_info", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)])
_cu_copy = _sig("xp_cuper_packed_copy", C.c_int, [_vp, _i32p, _u64pp])
_cu_create = _sig("xp_cuper_packed_create", C.c_int, [C.c_int, C.c_int, _i32p, _u64pp, C.POINTER(_vp)])
_cu_unpack = _sig("xp_cuper_unpack", C.c_int, [_vp, C.c_int, C.POINTER(_vp)])
_cu_free = _sig("xp_cuper_packed_free", None, [_vp])
_bm_pack = _sig("xp_bitmap_pack", C.c_int, [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.POINTER(_vp)])
_bm_info = _sig("xp_bitmap_info", C.c_int,
                [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)])
_bm_copy = _sig("xp_bitmap_copy", C.c_int, [_vp, _i32p, _i32p, _vp])
_bm_free = _sig("xp_bitmap_free", None, [_vp])
_bm_unpack = _sig("xp_bitmap_unpack", C.c_int,
                  [C.c_int, C.c_int, _i32p, _i32p, _vp, C.c_longlong, C.POINTER(_vp)])
_bm_spmv = _sig("xp_bitmap_spmv_f64", C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, C.c_longlong, _vp, _vp, _vp])
_bm_time = _sig("xp_bitmap_spmv_time", C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, C.c_longlong, _vp, _vp,
                                                   C.c_int, C.c_int, C.POINTER(C.c_double)])

class MtxInfo(C.Structure):
    _fields_ = [("nrows", C.c_longlong), ("ncols", C.c_longlong), ("nnz_reported", C.c_longlong),
                ("entries", C.c_longlong), ("field", C.c_int), ("symmetry", C.c_int), ("n", C.c_int),
                ("m", C.c_int), ("nnz", C.c_longlong), ("has_diag", C.c_int), ("diag_min_abs", C.c_double)]


_mtx_load = _sig("xp_mtx_load", C.c_int, [C.c_char_p, C.c_char_p, C.c_longlong, C.POINTER(MtxInfo), C.POINTER(_vp)])
_mtx_vector = _sig("xp_mtx_read_vector", C.c_int, [C.c_char_p, C.c_int, C.c_longlong, _f64p])
_ref_solution = _sig("xp_reference_solution", C.c_int, [C.c_int, _f64p])
_ds_save_bin = _sig("xp_dataset_save_binary", C.c_int, [_vp, C.c_char_p])
_ds_load_bin = _sig("xp_dataset_load_binary", C.c_int, [C.c_char_p, C.POINTER(_vp)])
_ds_load_cached = _sig("xp_dataset_load_cached", C.c_int,
                       [C.c_char_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(_vp)])
_sell_build = _sig("xp_sell_build_host", C.c_int,
                   [C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int,
                    C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_int)])
_sell_copy = _sig("xp_sell_copy_host", C.c_int, [_vp] * 9)
_m_create_dev = _sig("xp_matrix_create_device", C.c_int,
                     [C.c_int, _vp, _vp, _vp, C.POINTER(MatrixOpts), C.POINTER(_vp)])
_m_setup_stats = _sig("xp_matrix_setup_stats", C.c_int, [_vp, _i32p])
_m_create_bitmap = _sig("xp_matrix_create_bitmap", C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, C.c_longlong,
                                                             C.POINTER(MatrixOpts), C.POINTER(_vp)])
_m_csr_copy = _sig("xp_matrix_csr_copy", C.c_int, [_vp, _i32p, _i32p, _f64p])
_m_sell_copy = _sig("xp_matrix_sell_copy", C.c_int, [_vp] * 10)
_p_set_rhs_dev = _sig("xp_pcg_set_rhs_device", C.c_int, [_vp, _vp, _vp])
_p_get_x_dev = _sig("xp_pcg_get_x_device", C.c_int, [_vp, _vp])


class XplusError(RuntimeError):
    pass


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    _last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def _check(rc: int) -> None:
    if rc != 0:
        raise XplusError(last_error())


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- drop-in surface
def spmv_csr_kernel(row_ptr, col_idx, values, x, y, n: int) -> None:
    _spmv_csr(row_ptr, col_idx, values, x, y, n)


def spmv_blocked_kernel(row_ptr, col_idx, values, x, y, n: int) -> None:
    _spmv_blocked(row_ptr, col_idx, values, x, y, n)


def init_pcg_kernel(b, ax, m_inv, r, z, p, metrics, n: int) -> None:
    _init_pcg(b, ax, m_inv, r, z, p, metrics, n)


def dot_kernel(a, b, out, n: int) -> None:
    _dot(a, b, out, n)


def update_xrz_kernel(x, p, r, ap, m_inv, z, metrics, alpha: float, n: int) -> None:
    _update_xrz(x, p, r, ap, m_inv, z, metrics, alpha, n)


def update_p_kernel(z, p, beta: float, n: int) -> None:
    _update_p(z, p, beta, n)


def legacy_reset() -> None:
    _legacy_reset()


# ---------------------------------------------------------------- host CSR
class Csr:
    """Host CSR matrix with b and x0 (numpy), as CsrDataset holds them."""

    def __init__(self, row_ptr, col_idx, values, b=None, x0=None):
        self.row_ptr = _i32(row_ptr)
        self.col_idx = _i32(col_idx)
        self.values = _f64(values)
        self.n = len(self.row_ptr) - 1
        self.b = _f64(b) if b is not None else np.zeros(self.n)
        self.x0 = _f64(x0) if x0 is not None else np.zeros(self.n)

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    @classmethod
    def _from_handle(cls, h: _vp) -> "Csr":
        try:
            n, nnz = C.c_int(), C.c_int()
            _check(_csr_info(h, C.byref(n), C.byref(nnz)))
            rp = np.empty(n.value + 1, np.int32)
            ci = np.empty(max(nnz.value, 1), np.int32)
            va = np.empty(max(nnz.value, 1), np.float64)
            b = np.empty(max(n.value, 1), np.float64)
            x0 = np.empty(max(n.value, 1), np.float64)
            _check(_csr_copy(h, rp, ci, va, b, x0))
            return cls(rp, ci[:nnz.value], va[:nnz.value], b[:n.value], x0[:n.value])
        finally:
            _csr_free(h)

    def _handle(self) -> _vp:
        h = _vp()
        _check(_csr_from(self.n, self.row_ptr, self.col_idx, self.values, self.b, self.x0, C.byref(h)))
        return h

    def save(self, directory) -> None:
        h = self._handle()
        try:
            _check(_ds_save(h, str(directory).encode()))
        finally:
            _csr_free(h)

    def save_binary(self, path) -> None:
        """One-file binary copy (row_ptr, col_idx, values, b, x0 + checksum)."""
        h = self._handle()
        try:
            _check(_ds_save_bin(h, str(path).encode()))
        finally:
            _csr_free(h)

    def to_dense(self) -> np.ndarray:
        a = np.zeros((self.n, self.n))
        for r in range(self.n):
            for o in range(self.row_ptr[r], self.row_ptr[r + 1]):
                a[r, self.col_idx[o]] += self.values[o]
        return a

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.col_idx, self.row_ptr), shape=(self.n, self.n))


def load_dataset(directory) -> Csr:
    """CsrDataset::load semantics (src/CsrDataset.hpp:19-67); raises XplusError
    carrying the reference's message on invalid input."""
    h = _vp()
    _check(_ds_load(str(directory).encode(), C.byref(h)))
    return Csr._from_handle(h)


def load_binary(path) -> Csr:
    h = _vp()
    _check(_ds_load_bin(str(path).encode(), C.byref(h)))
    return Csr._from_handle(h)


def load_dataset_cached(directory, cache_path=None) -> tuple:
    """load_dataset through the binary cache: (Csr, True when the cache was
    used).  The cache (default directory/csr_cache.xpbin) is rewritten when
    the text files' sizes or modification times changed."""
    h = _vp()
    hit = C.c_int()
    cache = str(cache_path).encode() if cache_path is not None else None
    _check(_ds_load_cached(str(directory).encode(), cache, C.byref(hit), C.byref(h)))
    return Csr._from_handle(h), bool(hit.value)


MTX_FIELDS = ("real", "integer", "pattern")
MTX_SYMMETRIES = ("general", "symmetric")
X_REF_SOURCE = "A * deterministic smooth x_ref"


def load_mtx(path, b_path=None, subset_n=None) -> tuple:
    """MatrixMarket matrix -> (Csr, info), as the reference's converter
    (scripts/download_suitesparse_data.py:224-349,418-444) builds a dataset:
    b read from b_path, else b = A x_ref; x0 = 0; subset_n keeps the leading
    principal submatrix."""
    h = _vp()
    info = MtxInfo()
    bp = str(b_path).encode() if b_path else None
    _check(_mtx_load(str(path).encode(), bp, int(subset_n or 0), C.byref(info), C.byref(h)))
    d = {k: getattr(info, k) for k, _ in MtxInfo._fields_}
    d["field"] = MTX_FIELDS[info.field]
    d["symmetry"] = MTX_SYMMETRIES[info.symmetry]
    d["has_diag"] = bool(info.has_diag)
    d["diag_min_abs"] = info.diag_min_abs if info.has_diag else None
    return Csr._from_handle(h), d


def read_mtx_vector(path, expected_n: int, subset_n=None) -> np.ndarray:
    limit = min(expected_n, subset_n) if subset_n else expected_n
    out = np.zeros(max(limit, 1), np.float64)
    _check(_mtx_vector(str(path).encode(), expected_n, int(subset_n or 0), out))
    return out[:limit]


def reference_solution(n: int) -> np.ndarray:
    out = np.zeros(max(n, 1), np.float64)
    _check(_ref_solution(n, out))
    return out[:n]


def convert_mtx(mtx_path, out_dir, b_path=None, subset_n=None, name=None, source=None,
                description: str = "") -> dict:
    """convert_dataset (scripts/download_suitesparse_data.py:418-444) minus the
    download: writes row_ptr/col_idx/values/b/x0 .txt (%.17g) and
    meta.json/meta.txt (write_metadata, :379-415) into out_dir."""
    import json
    mtx_path, out_dir = Path(mtx_path), Path(out_dir)
    a, info = load_mtx(mtx_path, b_path, subset_n)
    out_dir.mkdir(parents=True, exist_ok=True)
    a.save(out_dir)
    name = name or out_dir.name
    meta = {"name": name, "source": source or str(mtx_path), "shape": [a.n, a.n], "nnz_csr": a.nnz,
            "matrix_market_path": str(mtx_path), "b_source": Path(b_path).name if b_path else X_REF_SOURCE,
            "x0": "zeros", "matrix_market_storage": "coordinate", "matrix_market_field": info["field"],
            "matrix_market_symmetry": info["symmetry"], "original_shape": [info["nrows"], info["ncols"]],
            "has_nonzero_diag_entries": info["has_diag"]}
    if description:
        meta["description"] = description
    if info["diag_min_abs"] is not None:
        meta["diag_min_abs"] = info["diag_min_abs"]
    if subset_n is not None:
        meta["subset_type"] = "leading_principal_submatrix"
        meta["subset_n"] = subset_n
    (out_dir / "meta.json").write_text(json.dumps(meta, indent=2, sort_keys=True) + "\n", encoding="utf-8")
    lines = [f"name={name}", f"source={meta['source']}", f"n={a.n}", f"m={a.n}", f"nnz={a.nnz}",
             f"b_source={meta['b_source']}", "x0=zeros"]
    if description:
        lines.append(f"description={description}")
    if subset_n is not None:
        lines += ["subset_type=leading_principal_submatrix", f"subset_n={subset_n}"]
    (out_dir / "meta.txt").write_text("\n".join(lines) + "\n", encoding="utf-8")
    return meta


class CuperPacked:
    """The reference's Cuper packed matrix (host/cuper_control_matrix.hpp):
    ptr = sp_element_list_ptr [batch_count + 1], data = uint64 [16, channel_size]
    (the 16 matrix_fpga_data channel buffers)."""

    def __init__(self, ptr, data):
        self.ptr = np.ascontiguousarray(ptr, np.int32)
        self.data = np.ascontiguousarray(data, np.uint64)
        self.batch_count = len(self.ptr) - 1
        self.channel_size = self.data.shape[1]

    @property
    def matrix_len(self) -> int:
        return int(self.ptr[-1])

    def _rows(self):
        rows = (C.POINTER(C.c_ulonglong) * 16)()
        for c in range(16):
            rows[c] = self.data[c].ctypes.data_as(C.POINTER(C.c_ulonglong))
        return rows

    def unpack(self, n: int) -> "Csr":
        h = _vp()
        _check(_cu_create(self.batch_count, self.channel_size, self.ptr, self._rows(), C.byref(h)))
        try:
            out = _vp()
            _check(_cu_unpack(h, n, C.byref(out)))
            return Csr._from_handle(out)
        finally:
            _cu_free(h)


# The reference's SpmvBlock (include/cg_common.hpp:47-50): 16 FP64 values,
# 16 index bytes whose bytes 2..3 hold the 16-bit occupancy mask.
SPMV_BLOCK = np.dtype([("values", "<f8", (16,)), ("indices", "u1", (16,))])
assert SPMV_BLOCK.itemsize == 144


class BitmapBlocks:
    """A matrix in the reference's 4x4 bitmap-block layout (host arrays):
    window_size 0 = plain block-row CSR (row_ptr of nbr + 1 entries), else
    window-major with num_windows * (nbr + 1) pointers."""

    def __init__(self, n: int, window_size: int, row_ptr, col_idx, blocks):
        self.n = int(n)
        self.window_size = int(window_size)
        self.row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, np.int32)
        self.blocks = np.ascontiguousarray(blocks, SPMV_BLOCK)
        # the C-ABI takes bare pointers: the array lengths are checked here
        want = self.num_windows * (self.num_block_rows + 1) if self.n > 0 else len(self.row_ptr)
        if len(self.row_ptr) != want:
            raise ValueError(f"row_ptr holds {len(self.row_ptr)} entries, the layout needs {want}")
        if len(self.col_idx) != len(self.blocks):
            raise ValueError("col_idx and blocks differ in length")

    @property
    def num_block_rows(self) -> int:
        return (self.n + 3) // 4

    @property
    def num_windows(self) -> int:
        return -(-self.n // self.window_size) if self.window_size > 0 else 1

    @property
    def nblocks(self) -> int:
        return len(self.blocks)

    def masks(self) -> np.ndarray:
        return self.blocks["indices"][:, 2].astype(np.uint32) | self.blocks["indices"][:, 3].astype(np.uint32) << 8

    def unpack(self) -> "Csr":
        """The CSR the layout describes (xp_bitmap_unpack)."""
        h = _vp()
        _check(_bm_unpack(self.n, self.window_size, self.row_ptr, self.col_idx, self.blocks.ctypes.data_as(_vp),
                          self.nblocks, C.byref(h)))
        return Csr._from_handle(h)

    def to_device(self) -> "DeviceBitmapBlocks":
        return DeviceBitmapBlocks(self)


def bitmap_pack(a: Csr, window_size: int = 512) -> BitmapBlocks:
    """build_windowed_blocked_spmv_matrix (archived/host/xrt_host.cpp:229-349);
    window_size 0: the plain layout."""
    h = _vp()
    _check(_bm_pack(a.n, a.row_ptr, a.col_idx, a.values, window_size, C.byref(h)))
    try:
        nw, nbr, nb = C.c_int(), C.c_int(), C.c_longlong()
        _check(_bm_info(h, C.byref(nw), C.byref(nbr), C.byref(nb)))
        rp = np.zeros(nw.value * (nbr.value + 1), np.int32)
        ci = np.zeros(nb.value, np.int32)
        blocks = np.zeros(nb.value, SPMV_BLOCK)
        _check(_bm_copy(h, rp, ci, blocks.ctypes.data_as(_vp)))
        return BitmapBlocks(a.n, window_size, rp, ci, blocks)
    finally:
        _bm_free(h)


class DeviceBitmapBlocks:
    """A bitmap-block layout uploaded to device memory."""

    def __init__(self, bm: BitmapBlocks):
        self.host = bm
        self.n = bm.n
        self.window_size = bm.window_size
        self.nblocks = bm.nblocks
        self.row_ptr = DeviceArray.from_numpy(bm.row_ptr if len(bm.row_ptr) else np.zeros(1, np.int32))
        self.col_idx = DeviceArray.from_numpy(bm.col_idx if bm.nblocks else np.zeros(1, np.int32))
        self.blocks = DeviceArray.from_numpy(bm.blocks.view(np.uint8).reshape(-1) if bm.nblocks
                                             else np.zeros(144, np.uint8))

    def _args(self):
        return (self.n, self.window_size, self.row_ptr.ptr, self.col_idx.ptr, self.blocks.ptr, self.nblocks)

    def spmv(self, x: "DeviceArray", y: "DeviceArray", stream: int = 0) -> None:
        """y = A x on device arrays of n doubles (xp_bitmap_spmv_f64)."""
        _check(_bm_spmv(*self._args(), x.ptr, y.ptr, stream or None))

    def spmv_numpy(self, x) -> np.ndarray:
        """y = A x for a host vector (uploads x, downloads y)."""
        xd = DeviceArray.from_numpy(_f64(x))
        yd = DeviceArray(max(self.n, 1), np.float64)
        try:
            self.spmv(xd, yd)
            synchronize()
            return yd.numpy()[:self.n]
        finally:
            xd.close()
            yd.close()

    def time(self, x: "DeviceArray", y: "DeviceArray", warmup: int = 10, reps: int = 100) -> float:
        """Average microseconds per SpMV launch (HIP events on its stream)."""
        us = C.c_double()
        _check(_bm_time(*self._args(), x.ptr, y.ptr, warmup, reps, C.byref(us)))
        return us.value

    def close(self) -> None:
        for k in ("row_ptr", "col_idx", "blocks"):
            getattr(self, k).close()


def cuper_pack(a: Csr) -> CuperPacked:
    h = _vp()
    _check(_cu_pack(a.n, a.row_ptr, a.col_idx, a.values, C.byref(h)))
    try:
        bc, ml, cs = C.c_int(), C.c_int(), C.c_int()
        _check(_cu_info(h, C.byref(bc), C.byref(ml), C.byref(cs)))
        ptr = np.zeros(bc.value + 1, np.int32)
        data = np.zeros((16, max(cs.value, 0)), np.uint64)
        p = CuperPacked(ptr, data)
        _check(_cu_copy(h, p.ptr, p._rows()))
        return p
    finally:
        _cu_free(h)


def gen_diffusion2d(nx: int, ny: int, sigma: float = 1.0, seed: int = 0) -> Csr:
    h = _vp()
    _check(_gen2d(nx, ny, sigma, seed, C.byref(h)))
    return Csr._from_handle(h)


def gen_diffusion3d(k: int, sigma: float = 1.5, seed: int = 0) -> Csr:
    h = _vp()
    _check(_gen3d(k, sigma, seed, C.byref(h)))
    return Csr._from_handle(h)


def gen_fem2d(m: int, sigma: float = 1.0, n_incl: int = 20, contrast: float = 1e3,
              jitter: float = 0.3, seed: int = 0) -> Csr:
    h = _vp()
    _check(_genfem(m, sigma, n_incl, contrast, jitter, seed, C.byref(h)))
    return Csr._from_handle(h)


def _sell_arrays(n_pad, nsl, stored, idx16):
    return {"perm": np.empty(n_pad, np.int32), "slice_ptr": np.empty(nsl + 1, np.int32),
            "slice_base": np.empty(nsl, np.int32), "val": np.empty(stored, np.float32),
            "val_lo": np.empty(stored, np.float32), "val_lo16": np.empty(stored, np.int16),
            "val64": np.empty(stored, np.float64),
            "col16": np.empty(stored if idx16 else 1, np.uint16), "col32": np.empty(1 if idx16 else stored, np.int32)}


def _ptrs(arrs, names):
    return [arrs[k].ctypes.data_as(_vp) if k in arrs else None for k in names]


_SELL_NAMES = ["perm", "slice_ptr", "slice_base", "val", "val_lo", "val_lo16", "val64", "col16", "col32"]


def sell_build_host(a: Csr, sigma: int = 256, keep_lo: bool = False, keep64: bool = False, block_rows: int = 1,
                    slots=None) -> dict:
    """Host-side SELL build (no GPU): returns the arrays of the format.
    block_rows = 8 with slots (units x 8 rows, -1 padding) keeps those
    blocks together; without slots, 8 consecutive rows."""
    n_pad, nsl, idx16 = C.c_int(), C.c_int(), C.c_int()
    stored = C.c_longlong()
    sl = None if slots is None else _i32(np.asarray(slots).ravel())
    units = 0 if sl is None else len(sl) // 8
    _check(_sell_build(a.n, a.row_ptr, a.col_idx, a.values, sigma, int(keep_lo), int(keep64), block_rows,
                       None if sl is None else sl.ctypes.data_as(_vp), units, C.byref(n_pad), C.byref(nsl),
                       C.byref(stored), C.byref(idx16)))
    arrs = _sell_arrays(n_pad.value, nsl.value, stored.value, idx16.value)
    if not keep_lo:
        del arrs["val_lo"], arrs["val_lo16"]
    if not keep64:
        del arrs["val64"]
    _check(_sell_copy(*_ptrs(arrs, _SELL_NAMES)))
    out = {"n_pad": n_pad.value, "nslices": nsl.value, "stored": stored.value, "idx16": idx16.value, **arrs}
    out["col16"] = arrs["col16"] if idx16.value else None
    out["col32"] = None if idx16.value else arrs["col32"]
    return out


# ---------------------------------------------------------------- device memory
class DeviceArray:
    """A device buffer owned through the library (hipMalloc), with numpy
    upload/download.  `ptr` is the device address for the xp_* calls."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        h = _vp()
        _check(_dev_alloc(self.count * self.dtype.itemsize, C.byref(h)))
        self.h = h

    @property
    def ptr(self) -> int:
        return self.h.value or 0

    @classmethod
    def from_numpy(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.count
        _check(_htod(self.h, a.ctypes.data_as(_vp), a.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        _check(_dtoh(out.ctypes.data_as(_vp), self.h, out.nbytes))
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        """Device-to-device copy of other's bytes (the runtime's copy kernel)."""
        n = min(self.count * self.dtype.itemsize, other.count * other.dtype.itemsize)
        _check(_dtod(self.h, other.h, n))

    def close(self) -> None:
        if getattr(self, "h", None) and self.h.value:
            _dev_free(self.h)
        self.h = None

    def __del__(self):
        self.close()


def stream_time(nbytes: int, mode: int = 0, grid: int = 2048, reps: int = 20) -> float:
    """Average us per launch of the streaming reference kernel over `nbytes`:
    read (mode 0), copy (1), non-temporal read (2), copy with non-temporal
    stores (3), copy with non-temporal loads and stores (4)."""
    us = C.c_double()
    _check(_stream_time(nbytes, mode, grid, reps, C.byref(us)))
    return us.value


class DeviceCsr:
    """A CSR system (and its b) uploaded once to device memory."""

    def __init__(self, a: Csr):
        self.host = a
        self.n = a.n
        self.row_ptr = DeviceArray.from_numpy(_i32(a.row_ptr))
        self.col_idx = DeviceArray.from_numpy(_i32(a.col_idx))
        self.values = DeviceArray.from_numpy(_f64(a.values))
        self.b = DeviceArray.from_numpy(_f64(a.b))

    def close(self) -> None:
        for k in ("row_ptr", "col_idx", "values", "b"):
            getattr(self, k).close()


def pool_trim() -> None:
    """Return the library's cached device blocks to the runtime."""
    _check(_pool_trim())


def pool_bytes() -> int:
    """Bytes of device memory the library's allocator holds for reuse."""
    v = C.c_ulonglong()
    _check(_pool_bytes(C.byref(v)))
    return v.value


def synchronize() -> None:
    _check(_dev_sync())


# ---------------------------------------------------------------- device matrix
class _Shape:
    """The size of a system whose CSR lives only on the device."""

    def __init__(self, n: int):
        self.n = int(n)


class Matrix:
    """Device-resident SELL-C-sigma matrix (one system or a batch)."""

    def __init__(self, systems, sigma: int = 0, keep_lo: bool = False, keep64: bool = False, grid: int = 0,
                 spmv_variant: int = 0, block_rows: int = 1, blocking: int = 0, device_build: bool = False,
                 _device_csr: "DeviceCsr" = None, _device_bitmap: "DeviceBitmapBlocks" = None):
        """blocking (with block_rows = 8): 0 blocks of 8 consecutive rows,
        1 greedy strength aggregates, 2 strength-ordered aggregates, 3 the
        parallel strength aggregation on the device, 4 the same with
        distance-1 seeds.  device_build: set the matrix up on the device (one
        system)."""
        if isinstance(systems, (Csr, _Shape)):
            systems = [systems]
        self.systems = list(systems)
        opts = MatrixOpts(sigma, int(keep_lo), int(keep64), grid, spmv_variant, block_rows, blocking,
                          int(device_build))
        h = _vp()
        if _device_bitmap is not None:
            d = _device_bitmap
            _check(_m_create_bitmap(*d._args(), C.byref(opts), C.byref(h)))
        elif _device_csr is not None:
            d = _device_csr
            _check(_m_create_dev(d.n, d.row_ptr.ptr, d.col_idx.ptr, d.values.ptr, C.byref(opts), C.byref(h)))
        elif len(self.systems) == 1:
            a = self.systems[0]
            _check(_m_create(a.n, a.row_ptr, a.col_idx, a.values, C.byref(opts), C.byref(h)))
        else:
            ns = _i32([a.n for a in self.systems])
            rp = _i32(np.concatenate([a.row_ptr for a in self.systems]))
            ci = _i32(np.concatenate([a.col_idx for a in self.systems]))
            va = _f64(np.concatenate([a.values for a in self.systems]))
            _check(_m_create_batch(len(self.systems), ns, rp, ci, va, C.byref(opts), C.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            _m_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @classmethod
    def from_device(cls, d: "DeviceCsr", **kw) -> "Matrix":
        """A matrix set up on the device from a CSR already in device memory."""
        return cls(d.host, _device_csr=d, **kw)

    @classmethod
    def from_bitmap(cls, d: "DeviceBitmapBlocks", **kw) -> "Matrix":
        """A matrix set up on the device from a bitmap-block layout in device
        memory (xp_matrix_create_bitmap)."""
        return cls(_Shape(d.n), _device_bitmap=d, **kw)

    def csr(self):
        """The CSR the matrix was built from: (row_ptr, col_idx, values)."""
        i = self.info()
        rp = np.empty(i["n"] + 1, np.int32)
        ci = np.empty(max(i["nnz"], 1), np.int32)
        va = np.empty(max(i["nnz"], 1), np.float64)
        _check(_m_csr_copy(self.h, rp, ci, va))
        return rp, ci[:i["nnz"]], va[:i["nnz"]]

    def setup_stats(self) -> dict:
        st = np.zeros(5, np.int32)
        _check(_m_setup_stats(self.h, st))
        return {"device_built": bool(st[0]), "rounds": int(st[1]), "regroup_rounds": int(st[2]),
                "pieces": int(st[3]), "blocks": int(st[4])}

    def sell_arrays(self) -> dict:
        """The device SELL arrays (downloaded)."""
        i = self.info()
        arrs = _sell_arrays(i["n_pad"], i["nslices"], i["stored"], i["idx16"])
        _check(_m_sell_copy(self.h, *_ptrs(arrs, _SELL_NAMES)))
        return {**i, **arrs, "col16": arrs["col16"] if i["idx16"] else None,
                "col32": None if i["idx16"] else arrs["col32"]}

    def num_blocks(self) -> int:
        return _m_num_blocks(self.h)

    def block_rows(self) -> np.ndarray:
        """Rows of every block (global natural rows, -1 padding): nblocks x 8."""
        out = np.empty(self.num_blocks() * 8, np.int32)
        _check(_m_block_rows(self.h, out))
        return out.reshape(-1, 8)

    def block_features(self):
        """Device-computed block features (natural block order): feat nb x 48, scale nb x 8."""
        nb = self.num_blocks()
        f = np.empty(nb * 48, np.float32)
        sc = np.empty(nb * 8, np.float32)
        _check(_m_block_features(self.h, f, sc))
        return f.reshape(nb, 48), sc.reshape(nb, 8)

    def info(self) -> dict:
        i = MatrixInfo()
        _check(_m_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def permutation(self) -> np.ndarray:
        perm = np.empty(self.info()["n_pad"], np.int32)
        _check(_m_perm(self.h, perm))
        return perm

    def set_grid(self, grid: int) -> None:
        _check(_m_set_grid(self.h, grid))

    def set_variant(self, v: int) -> None:
        _check(_m_set_variant(self.h, v))

    def autotune(self, fused=False, tune_sigma: bool = False, iters: int = 64) -> dict:
        """Picks SpMV inner loop x workgroup count (x sigma) by timing real PCG
        iterations (fused: False/True, or 2 for the accurate iteration);
        tune_sigma re-sorts the matrix, so call it before any Pcg."""
        r = TuneReport()
        _check(_m_autotune(self.h, int(fused), int(tune_sigma), iters, C.byref(r)))
        return {k: getattr(r, k) for k, _ in r._fields_}

    # device-pointer operations: ints are device addresses in the library's HIP
    # runtime (DeviceArray.ptr, or hipMalloc through the same libamdhip64)
    def spmv_f32(self, x_ptr: int, y_ptr: int, stream: int = 0) -> None:
        _check(_spmv_f32(self.h, x_ptr, y_ptr, stream or None))

    def spmv_f64(self, x_ptr: int, y_ptr: int, stream: int = 0) -> None:
        _check(_spmv_f64(self.h, x_ptr, y_ptr, stream or None))

    # the fused vector operations of a PCG iteration on solve-order FP32
    # vectors (n_pad entries); metrics / out are FP64 device arrays:
    # [2 k] = r^T z, [2 k + 1] = r^T r per system k, out[k] = a^T b
    def init_pcg_f32(self, b_ptr, ax_ptr, minv_ptr, r_ptr, z_ptr, p_ptr, metrics_ptr, stream: int = 0) -> None:
        _check(_op_init(self.h, b_ptr, ax_ptr, minv_ptr, r_ptr, z_ptr, p_ptr, metrics_ptr, stream or None))

    def dot_f32(self, a_ptr, b_ptr, out_ptr, stream: int = 0) -> None:
        _check(_op_dot(self.h, a_ptr, b_ptr, out_ptr, stream or None))

    def update_xrz_f32(self, x_ptr, p_ptr, r_ptr, ap_ptr, minv_ptr, z_ptr, metrics_ptr, alpha: float,
                       stream: int = 0) -> None:
        _check(_op_xrz(self.h, x_ptr, p_ptr, r_ptr, ap_ptr, minv_ptr, z_ptr, metrics_ptr, float(alpha),
                       stream or None))

    def update_p_f32(self, z_ptr, p_ptr, beta: float, stream: int = 0) -> None:
        _check(_op_p(self.h, z_ptr, p_ptr, float(beta), stream or None))

    def spmv_f32_time(self, warmup: int = 20, reps: int = 200) -> float:
        """Mean microseconds per FP32 SpMV launch (HIP events on the launch stream)."""
        us = C.c_double()
        _check(_spmv_time(self.h, warmup, reps, C.byref(us)))
        return us.value

    def spmv_f32_time_cold(self, reps: int = 20, flush_bytes: int = 1 << 30) -> float:
        """The same with the caches flushed (a flush_bytes copy) before every launch."""
        us = C.c_double()
        _check(_spmv_time_cold(self.h, reps, flush_bytes, C.byref(us)))
        return us.value

    def to_solve_order(self, nat_ptr: int, solve_ptr: int, stream: int = 0) -> None:
        _check(_to_solve(self.h, nat_ptr, solve_ptr, stream or None))

    def from_solve_order(self, solve_ptr: int, nat_ptr: int, stream: int = 0) -> None:
        _check(_from_solve(self.h, solve_ptr, nat_ptr, stream or None))


def csr_spmv_f64(a: Csr, x) -> np.ndarray:
    y = np.zeros(a.n)
    _check(_csr_spmv_f64(a.n, a.row_ptr, a.col_idx, a.values, _f64(x), y))
    return y


# ---------------------------------------------------------------- solver
class Jacobi:
    """Plain Jacobi iteration x_new = D^-1 (b - R x) in FP32 over a Matrix
    (the reference's Cuper-jacobi-iteration host and kernel contract):
    max_iters updates, or fewer with use_tau once max |x_new - x_old| <= tau."""

    def __init__(self, matrix: "Matrix", max_iters: int = 1, tau: float = 1e-5, use_tau: bool = False,
                 graph_iters: int = 0, use_graph: bool = True):
        self.matrix = matrix
        self.opts = JacobiOpts(tau, max_iters, int(use_tau), graph_iters, int(use_graph))
        h = _vp()
        _check(_j_create(matrix.h, C.byref(self.opts), C.byref(h)))
        self.h = h
        self.nsys = len(matrix.systems)
        self.n = sum(a.n for a in matrix.systems)

    def close(self) -> None:
        if getattr(self, "h", None):
            _j_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_rhs(self, b, x0=None) -> None:
        b = _f64(b)
        if x0 is None:
            _check(_j_set_rhs(self.h, b, None))
        else:
            x0 = _f64(x0)
            _check(_j_set_rhs(self.h, b, x0.ctypes.data_as(_vp)))

    def run(self) -> None:
        _check(_j_run(self.h))

    def status(self, sys: int = 0) -> dict:
        s = JacobiStatus()
        _check(_j_status(self.h, sys, C.byref(s)))
        return {"status": s.status, "iterations": s.iterations, "diff": s.diff}

    def x(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1))
        _check(_j_get_x(self.h, out))
        return out[:self.n]

    def timing(self) -> tuple:
        ms, launches = C.c_double(), C.c_int()
        _check(_j_timing(self.h, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value

    def kernel_time(self, iters: int = 20) -> float:
        us = C.c_double()
        _check(_j_kernel_time(self.h, iters, C.byref(us)))
        return us.value


def rhs_ones_f32(a: Csr) -> np.ndarray:
    """A * ones in FP32, row by row in CSR order: the reference's right-hand
    side for a dataset without b (Cuper-jacobi-iteration host/main.cpp:969-973)."""
    out = np.zeros(max(a.n, 1))
    _check(_rhs_ones(a.n, a.row_ptr, a.col_idx, a.values, out))
    return out[:a.n]


class Pcg:
    """Graph-captured Jacobi-PCG over a Matrix (one or many systems)."""

    def __init__(self, matrix: Matrix, tau: float = 1e-10, rel_tol: float = 0.0, max_iters: int = 0,
                 graph_iters: int = 0, use_graph: bool = True, trace_cap: int = 0, stall_limit: int = 0,
                 fused: bool = False, accurate: bool = False, full_work_map: bool = False):
        """tau: absolute bound on r^T r; rel_tol > 0 overrides it per system as
        (rel_tol ||b_k||)^2.  accurate: FP64-accumulated SpMV (hi+lo values,
        the matrix needs keep_lo) with FP64 p, A p and x."""
        self.matrix = matrix
        self.opts = SolveOpts(tau, rel_tol, max_iters, graph_iters, int(use_graph), trace_cap, stall_limit,
                              int(fused), int(accurate), int(full_work_map))  # accurate: 1 (hi + lo16) or 2 (hi + lo)
        h = _vp()
        _check(_p_create(matrix.h, C.byref(self.opts), C.byref(h)))
        self.h = h
        self.nsys = len(matrix.systems)
        self.n = sum(a.n for a in matrix.systems)

    def close(self) -> None:
        if getattr(self, "h", None):
            _p_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_rhs(self, b, x0=None) -> None:
        b = _f64(b)
        if x0 is None:
            _check(_p_set_rhs(self.h, b, None))
        else:
            x0 = _f64(x0)
            _check(_p_set_rhs(self.h, b, x0.ctypes.data_as(_vp)))

    def set_rhs_device(self, b_ptr: int, x0_ptr: int = 0) -> None:
        """b (and x0) from device memory: natural order, FP64."""
        _check(_p_set_rhs_dev(self.h, b_ptr, x0_ptr or None))

    def x_device(self, out_ptr: int) -> None:
        _check(_p_get_x_dev(self.h, out_ptr))

    def run(self) -> None:
        _check(_p_run(self.h))

    def run_fixed(self, iters: int) -> None:
        _check(_p_run_fixed(self.h, iters))

    def run_refined(self, outer_rel: float = 1e-10, inner_rel: float = 1e-5, max_outer: int = 20):
        rel = np.zeros(self.nsys)
        outer, total = C.c_int(), C.c_int()
        _check(_p_run_refined(self.h, outer_rel, inner_rel, max_outer, rel, C.byref(outer), C.byref(total)))
        return rel, outer.value, total.value

    def run_reliable(self, rel_tol: float = 1e-6, delta: float = 0.1):
        """FP32 PCG with residual replacement to a true relative residual:
        returns (true relative residuals, replacements, iterations)."""
        rel = np.zeros(self.nsys)
        repl, total = C.c_int(), C.c_int()
        _check(_p_run_reliable(self.h, rel_tol, delta, rel, C.byref(repl), C.byref(total)))
        return rel, repl.value, total.value

    def status(self, sys: int = 0) -> dict:
        s = Status()
        _check(_p_status(self.h, sys, C.byref(s)))
        return s.as_dict()

    def x(self) -> np.ndarray:
        out = np.zeros(self.n)
        _check(_p_get_x(self.h, out))
        return out

    def trace(self, sys: int = 0) -> np.ndarray:
        """Rows {alpha, beta, rz, rr} per completed iteration."""
        cap = max(self.opts.trace_cap, 1)
        rows = np.zeros(cap * 4)
        got = _p_trace(self.h, sys, rows, cap)
        if got < 0:
            raise XplusError(last_error())
        return rows[: got * 4].reshape(got, 4)

    def true_residual(self) -> np.ndarray:
        out = np.zeros(self.nsys)
        _check(_p_true_res(self.h, out))
        return out

    def set_preconditioner(self, mode: int, blocks=None) -> None:
        """0 Jacobi, 1 exact 8x8 block Jacobi, 2 given blocks (nblocks x 8 x 8, natural order)."""
        if blocks is None:
            _check(_p_set_precond(self.h, mode, None, 0))
        else:
            b = np.ascontiguousarray(blocks, dtype=np.float32)
            _check(_p_set_precond(self.h, mode, b.ctypes.data_as(_vp), b.size))

    def block_operator(self) -> np.ndarray:
        """The blocks the device applies (S C C^T S of the stored FP16
        factors, FP64), block order: nblocks x 8 x 8."""
        nb = self.matrix.num_blocks()
        out = np.zeros(nb * 64)
        _check(_p_block_op(self.h, out))
        return out.reshape(nb, 8, 8)

    def set_mlp_preconditioner(self, net: "Mlp") -> None:
        _check(_p_set_mlp(self.h, net.h))

    def set_multilevel(self, coarse_max: int = 512, omega: float = 0.8, gamma: float = 1.6,
                       coarse_blocking: int = 1, consecutive_rows: int = 0, theta: float = 0.08) -> None:
        """Multilevel aggregation preconditioner with the installed blocks as
        the level-0 smoother (coarse_max <= 0: off): coarse spaces from plain
        aggregation at strength threshold theta; the coarse levels' smoother
        blocks are pieces of their own aggregates (coarse_blocking 1) or their
        strength aggregates (3 or 4), those of at most consecutive_rows rows
        take consecutive rows."""
        o = MultilevelOpts(coarse_max, omega, gamma, coarse_blocking, consecutive_rows, theta)
        _check(_p_set_ml(self.h, C.byref(o)))

    def multilevel_info(self) -> dict:
        """Levels (rows, stored entries; the last is the dense coarsest) and set-up ms."""
        cap = 16
        nl, ms = C.c_int(), C.c_double()
        n, nnz = np.zeros(cap, np.int32), np.zeros(cap, np.int64)
        _check(_p_ml_info(self.h, cap, C.byref(nl), n, nnz, C.byref(ms)))
        k = min(nl.value, cap)
        ph, cnt = np.zeros(3 * cap), C.c_int()
        _check(_p_ml_phases(self.h, 3 * cap, ph, C.byref(cnt)))
        return {"levels": nl.value, "n": n[:k].tolist(), "nnz": nnz[:k].tolist(), "setup_ms": ms.value,
                "phase_ms": [round(v, 3) for v in ph[:min(cnt.value, 3 * cap)]]}

    def multilevel_level(self, level: int):
        """Level `level` >= 1 as a natural-order Csr (rows = the aggregates of
        the level above) and the aggregate of every row of the level above."""
        info = self.multilevel_info()
        n, nnz = info["n"][level], info["nnz"][level]
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        va = np.zeros(max(nnz, 1), np.float64)
        agg = np.zeros(max(info["n"][level - 1], 1), np.int32)
        _check(_p_ml_level(self.h, level, rp.ctypes.data_as(_vp), ci.ctypes.data_as(_vp), va.ctypes.data_as(_vp),
                           agg.ctypes.data_as(_vp)))
        return Csr(rp, ci[:nnz], va[:nnz]), agg[:info["n"][level - 1]]

    def multilevel_slots(self, level: int) -> np.ndarray:
        """The smoother blocks of coarse SELL level `level` (units x 8 rows of
        that level, -1 padding)."""
        units = C.c_int()
        _check(_p_ml_slots(self.h, level, None, C.byref(units)))
        slots = np.zeros(max(units.value, 1) * 8, np.int32)
        _check(_p_ml_slots(self.h, level, slots.ctypes.data_as(_vp), C.byref(units)))
        return slots[:units.value * 8].reshape(-1, 8)

    def multilevel_rounds(self) -> list:
        """Root-selection rounds of each level's aggregation."""
        r, cnt = np.zeros(16, np.int32), C.c_int()
        _check(_p_ml_rounds(self.h, 16, r, C.byref(cnt)))
        return r[:min(cnt.value, 16)].tolist()

    def multilevel_time(self, reps: int = 50) -> float:
        """Average device us of one V-cycle (block product included)."""
        us = C.c_double()
        _check(_p_ml_time(self.h, reps, C.byref(us)))
        return us.value

    def multilevel_apply(self, r_solve) -> np.ndarray:
        """z = M r then one V-cycle, r in level-0 solve order (n_pad FP32)."""
        r = np.ascontiguousarray(r_solve, dtype=np.float32)
        dr = DeviceArray.from_numpy(r)
        dz = DeviceArray(r.size, np.float32)
        try:
            _check(_p_ml_apply(self.h, dr.ptr, dz.ptr))
            return dz.numpy()
        finally:
            dr.close()
            dz.close()

    def kernel_times(self, iters: int = 100):
        """Average device time (us) of each kernel of the iteration (HIP events)."""
        us = np.zeros(3)
        nk = C.c_int()
        _check(_p_kernel_times(self.h, iters, us, C.byref(nk)))
        return us[:nk.value]

    def kernel_times_split(self, iters: int = 100):
        """kernel_times split by iteration parity: (even, odd) averages."""
        us = np.zeros(6)
        nk = C.c_int()
        _check(_p_kernel_times_split(self.h, iters, us, C.byref(nk)))
        return us[:nk.value], us[3:3 + nk.value]

    def last_timing(self):
        ms, launches = C.c_double(), C.c_int()
        _check(_p_timing(self.h, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value


# ---------------------------------------------------------------- learned block preconditioner
class BlockSet:
    """Features / training targets of the aligned 8-row blocks (host)."""

    def __init__(self, a: Csr = None, blocking: int = 0, lam: float = -1.0, slots=None):
        """lam: off-block weight of the training loss (default 0.5).
        blocking 3 (the device aggregation) computes the blocks on the GPU;
        slots (nblocks x 8, -1 padding) gives them explicitly."""
        self.h = None
        if a is not None:
            h = _vp()
            if blocking == 3 and slots is None:
                m = Matrix(a, block_rows=8, blocking=3)
                slots = m.block_rows()
                m.close()
            if slots is not None:
                sl = _i32(np.asarray(slots).ravel())
                _check(_bs_from_slots(a.n, a.row_ptr, a.col_idx, a.values, sl, len(sl) // 8, lam, C.byref(h)))
            else:
                _check(_bs_from_blocked(a.n, a.row_ptr, a.col_idx, a.values, blocking, lam, C.byref(h)))
            self.h = h

    def append(self, other: "BlockSet") -> None:
        _check(_bs_append(self.h, other.h))

    def __len__(self) -> int:
        return _bs_size(self.h)

    def arrays(self):
        n = len(self)
        f = np.empty(n * 48, np.float32)
        a = np.empty(n * 64, np.float32)
        hm = np.empty(n * 64, np.float32)
        sc = np.empty(n * 8, np.float32)
        _check(_bs_copy(self.h, f, a, hm, sc))
        return f.reshape(n, 48), a.reshape(n, 8, 8), hm.reshape(n, 8, 8), sc.reshape(n, 8)

    def close(self):
        if getattr(self, "h", None):
            _bs_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def blockset(systems, blocking: int = 0) -> BlockSet:
    if isinstance(systems, Csr):
        systems = [systems]
    out = BlockSet(systems[0], blocking)
    for a in systems[1:]:
        out.append(BlockSet(a, blocking))
    return out


def strength_blocking(a: Csr, rule: int = 1) -> np.ndarray:
    """Greedy strength aggregation of a (no GPU): nblocks x 8 local rows, -1
    padding.  rule 1: seeds in row order, growth by summed strength; rule 2:
    seeds by decreasing total strength, growth by the fraction of the
    candidate's strength the block holds."""
    slots = np.empty(8 * max(a.n, 1), np.int32)
    nb = C.c_int()
    if rule == 1:
        _check(_strength_blocking(a.n, a.row_ptr, a.col_idx, a.values, slots, C.byref(nb)))
    else:
        _check(_aggregate_blocking(a.n, a.row_ptr, a.col_idx, a.values, rule, slots, C.byref(nb)))
    return slots[:8 * nb.value].reshape(-1, 8)


class Mlp:
    """The block-preconditioner network on the GPU (BF16 MFMA forward)."""

    def __init__(self, seed: int = 0):
        h = _vp()
        _check(_mlp_create(seed, C.byref(h)))
        self.h = h

    @staticmethod
    def num_params() -> int:
        return _mlp_nparams()

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params(), np.float32)
        _check(_mlp_get(self.h, out))
        return out

    def set_params(self, p) -> None:
        _check(_mlp_set(self.h, np.ascontiguousarray(p, dtype=np.float32)))

    def loss_grad(self, bs: BlockSet, first: int = 0, count: int = None):
        count = len(bs) - first if count is None else count
        loss = C.c_double()
        g = np.empty(self.num_params(), np.float32)
        _check(_mlp_loss_grad(self.h, bs.h, first, count, C.byref(loss), g))
        return loss.value, g

    def forward(self, bs: BlockSet, first: int = 0, count: int = None) -> np.ndarray:
        count = len(bs) - first if count is None else count
        out = np.empty(count * 48, np.float32)
        _check(_mlp_forward(self.h, bs.h, first, count, out))
        return out.reshape(count, 48)

    def blocks(self, bs: BlockSet) -> np.ndarray:
        out = np.empty(len(bs) * 64, np.float32)
        _check(_mlp_blocks(self.h, bs.h, out))
        return out.reshape(len(bs), 8, 8)

    def train(self, bs: BlockSet, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        hist = np.zeros(max(steps, 1))
        _check(_mlp_train(self.h, bs.h, steps, batch, lr, seed, hist))
        return hist[:steps]

    def close(self):
        if getattr(self, "h", None):
            _mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
