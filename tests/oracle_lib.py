"""ctypes access to the checkers: oracle/liboracle.so (our C restatement of the
reference algorithm) and, when it was built in this container,
oracle/_ref/libref.so (the reference's own sources).  Test infrastructure:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline only.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libref.so"
REF_BIN = ROOT / "oracle" / "_ref" / "xplus_local"
REF_ON_AMD_BIN = ROOT / "oracle" / "_ref" / "xplus_local_on_amd"
REF_CUPER_ON_AMD_BIN = ROOT / "oracle" / "_ref" / "cuper_pcg_on_amd"
REF_BLOCKED_SO = ROOT / "oracle" / "_ref" / "libref_blocked.so"

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class Trace(C.Structure):
    _fields_ = [("iteration", C.c_int), ("alpha", C.c_double), ("beta", C.c_double),
                ("rz", C.c_double), ("rr", C.c_double), ("residual", C.c_double)]


class Dataset(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("row_ptr", C.POINTER(C.c_int)),
                ("col_idx", C.POINTER(C.c_int)), ("values", C.POINTER(C.c_double)),
                ("b", C.POINTER(C.c_double)), ("x0", C.POINTER(C.c_double))]


def _load_oracle():
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "all"], check=True)
    lib = C.CDLL(str(ORACLE_SO))
    lib.orc_dataset_load.argtypes = [C.c_char_p, C.POINTER(C.POINTER(Dataset)), C.c_char_p, C.c_int]
    lib.orc_dataset_free.argtypes = [C.POINTER(Dataset)]
    lib.orc_spmv_csr.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
    lib.orc_extract_jacobi_diag.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_char_p, C.c_int]
    lib.orc_run_jacobi_pcg.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                                       _f64p, C.POINTER(Trace), C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                       C.c_char_p, C.c_int]
    lib.orc_compute_residual_norm.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
    lib.orc_compute_residual_norm.restype = C.c_double
    lib.orc_k_spmv_csr.argtypes = [_i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]
    lib.orc_k_init_pcg.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int]
    lib.orc_k_dot.argtypes = [_f64p, _f64p, _f64p, C.c_int]
    lib.orc_k_update_xrz.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int]
    lib.orc_k_update_p.argtypes = [_f64p, _f64p, C.c_double, C.c_int]
    lib.orc_multi_kernel_solver.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                                            _f64p, C.POINTER(Trace), C.c_int, C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_int), _f64p, C.c_char_p, C.c_int]
    lib.orc_cuper_style_spmv.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
    _f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    lib.orc_jacobi_prepare_f32.argtypes = [C.c_int, _i32p, _i32p, _f64p, _vp, _f32p, _i32p, _i32p, _f32p, _f32p]
    lib.orc_jacobi_run_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_double,
                                       _f32p, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    return lib


oracle = _load_oracle()
ref = None
if REF_SO.exists():
    ref = C.CDLL(str(REF_SO))
    ref.ref_dataset_info.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int]
    ref.ref_dataset_read.argtypes = [C.c_char_p, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_char_p, C.c_int]
    ref.ref_dataset_spmv.argtypes = [C.c_char_p, _f64p, _f64p, C.c_char_p, C.c_int]
    ref.ref_extract_jacobi_diag.argtypes = [C.c_char_p, _f64p, C.c_char_p, C.c_int]
    ref.ref_golden.argtypes = [C.c_char_p, C.c_double, C.c_int, _f64p, _f64p, C.c_int, C.POINTER(C.c_int),
                               C.POINTER(C.c_int), _f64p, C.c_char_p, C.c_int]
    ref.ref_multi_kernel.argtypes = [C.c_char_p, C.c_double, C.c_int, _f64p, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), _f64p, C.c_char_p, C.c_int]
    ref.ref_cuper_style_spmv.argtypes = [C.c_char_p, _f64p, _f64p, C.c_char_p, C.c_int]
    if hasattr(ref, "ref_cuper_pack_info"):
        ref.ref_cuper_pack_info.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.c_char_p, C.c_int]
        ref.ref_cuper_pack_copy.argtypes = [C.c_char_p, _i32p,
                                            np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS"),
                                            C.c_char_p, C.c_int]
    for name in ("spmv_csr_kernel", "spmv_blocked_kernel"):
        getattr(ref, name).argtypes = [_i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]
    ref.init_pcg_kernel.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int]
    ref.dot_kernel.argtypes = [_f64p, _f64p, _f64p, C.c_int]
    ref.update_xrz_kernel.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_int]
    ref.update_p_kernel.argtypes = [_f64p, _f64p, C.c_double, C.c_int]


class OracleError(RuntimeError):
    pass


def load(directory):
    """orc_dataset_load -> dict of numpy arrays; raises OracleError(msg)."""
    d = C.POINTER(Dataset)()
    err = C.create_string_buffer(4096)
    if oracle.orc_dataset_load(str(directory).encode(), C.byref(d), err, 4096) != 0:
        raise OracleError(err.value.decode())
    try:
        n, nnz = d.contents.n, d.contents.nnz
        out = {
            "n": n,
            "row_ptr": np.ctypeslib.as_array(d.contents.row_ptr, (n + 1,)).copy(),
            "col_idx": np.ctypeslib.as_array(d.contents.col_idx, (max(nnz, 1),))[:nnz].copy(),
            "values": np.ctypeslib.as_array(d.contents.values, (max(nnz, 1),))[:nnz].copy(),
            "b": np.ctypeslib.as_array(d.contents.b, (n,)).copy(),
            "x0": np.ctypeslib.as_array(d.contents.x0, (n,)).copy(),
        }
    finally:
        oracle.orc_dataset_free(d)
    return out


def spmv(a, x):
    y = np.zeros(a.n)
    oracle.orc_spmv_csr(a.n, a.row_ptr, a.col_idx, a.values, np.ascontiguousarray(x, np.float64), y)
    return y


def golden(a, b=None, x0=None, tau=1e-10, max_iters=0, trace_cap=None):
    """orc_run_jacobi_pcg: dict(x, trace[iteration, alpha, beta, rz, rr, residual], iterations, final_rr,
    converged).  Raises OracleError with the reference's message."""
    b = np.ascontiguousarray(a.b if b is None else b, np.float64)
    x0 = np.ascontiguousarray(a.x0 if x0 is None else x0, np.float64)
    cap = trace_cap if trace_cap is not None else (max_iters if max_iters > 0 else max(4 * a.n, 1000))
    tr = (Trace * max(cap, 1))()
    x = np.zeros(a.n)
    tl, it, cv = C.c_int(), C.c_int(), C.c_int()
    frr = C.c_double()
    err = C.create_string_buffer(512)
    rc = oracle.orc_run_jacobi_pcg(a.n, a.row_ptr, a.col_idx, a.values, b, x0, tau, max_iters, x, tr, cap,
                                   C.byref(tl), C.byref(it), C.byref(frr), C.byref(cv), err, 512)
    if rc != 0:
        raise OracleError(err.value.decode())
    m = min(tl.value, cap)
    trace = np.array([[tr[i].iteration, tr[i].alpha, tr[i].beta, tr[i].rz, tr[i].rr, tr[i].residual]
                      for i in range(m)]).reshape(m, 6)
    return {"x": x, "trace": trace, "iterations": it.value, "final_rr": frr.value, "converged": bool(cv.value)}


def residual_norm(a, x, b=None):
    b = np.ascontiguousarray(a.b if b is None else b, np.float64)
    return oracle.orc_compute_residual_norm(a.n, a.row_ptr, a.col_idx, a.values, b,
                                            np.ascontiguousarray(x, np.float64))


def multi_kernel(a, tau=1e-10, max_iters=0, b=None, x0=None):
    b = np.ascontiguousarray(a.b if b is None else b, np.float64)
    x0 = np.ascontiguousarray(a.x0 if x0 is None else x0, np.float64)
    x = np.zeros(a.n)
    it, stt, emi = C.c_int(), C.c_int(), C.c_int()
    sc = np.zeros(3)
    err = C.create_string_buffer(512)
    rc = oracle.orc_multi_kernel_solver(a.n, a.row_ptr, a.col_idx, a.values, b, x0, tau, max_iters, x, None, 0,
                                        C.byref(it), C.byref(stt), C.byref(emi), sc, err, 512)
    if rc != 0:
        raise OracleError(err.value.decode())
    return {"x": x, "iterations": it.value, "status": stt.value, "effective_max_iters": emi.value,
            "final_rr": sc[0], "residual_l2": sc[1], "residual_rel": sc[2]}


def jacobi_prepare(a, b=None):
    """orc_jacobi_prepare_f32: dict(b, dinv, r_rp, r_ci, r_val) in FP32; b=None
    builds A * ones.  Raises OracleError('missing or zero diagonal at row R')."""
    n = a.n
    bb = np.zeros(max(n, 1), np.float32)
    dinv = np.zeros(max(n, 1), np.float32)
    rrp = np.zeros(n + 1, np.int32)
    rci = np.zeros(max(a.nnz, 1), np.int32)
    rva = np.zeros(max(a.nnz, 1), np.float32)
    bp = None if b is None else np.ascontiguousarray(b, np.float64)
    bad = oracle.orc_jacobi_prepare_f32(n, a.row_ptr, a.col_idx, a.values,
                                        None if bp is None else bp.ctypes.data_as(_vp), bb, rrp, rci, rva, dinv)
    if bad >= 0:
        raise OracleError(f"missing or zero diagonal at row {bad}")
    return {"b": bb[:n], "dinv": dinv[:n], "r_rp": rrp, "r_ci": rci, "r_val": rva}


def jacobi(a, b=None, x0=None, max_iters=1, use_tau=False, tau=1e-5):
    """orc_jacobi_run_f32 after jacobi_prepare: dict(x (float32), iterations,
    diff, status, b, dinv)."""
    pr = jacobi_prepare(a, b)
    x = np.zeros(max(a.n, 1), np.float32)
    if x0 is not None:
        x[:a.n] = np.asarray(x0, np.float64).astype(np.float32)
    fd, st = C.c_float(), C.c_int()
    bb = np.ascontiguousarray(np.r_[pr["b"], np.zeros(1, np.float32)])
    dv = np.ascontiguousarray(np.r_[pr["dinv"], np.zeros(1, np.float32)])
    it = oracle.orc_jacobi_run_f32(a.n, pr["r_rp"], pr["r_ci"], pr["r_val"], bb, dv, max_iters, int(use_tau), tau,
                                   x, C.byref(fd), C.byref(st))
    return {"x": x[:a.n], "iterations": it, "diff": fd.value, "status": st.value, "b": pr["b"], "dinv": pr["dinv"]}


def ref_cuper_pack(directory):
    """The reference's own build_cuper_control_matrix over a dataset
    directory: (ptr int32 [batches + 1], data uint64 [16, channel_size])."""
    bc, ml, cs = C.c_int(), C.c_int(), C.c_int()
    err = C.create_string_buffer(512)
    if ref.ref_cuper_pack_info(str(directory).encode(), C.byref(bc), C.byref(ml), C.byref(cs), err, 512) != 0:
        raise OracleError(err.value.decode())
    ptr = np.zeros(bc.value + 1, np.int32)
    data = np.zeros((16, max(cs.value, 1)), np.uint64)
    if ref.ref_cuper_pack_copy(str(directory).encode(), ptr, data, err, 512) != 0:
        raise OracleError(err.value.decode())
    return ptr, data[:, :cs.value]
