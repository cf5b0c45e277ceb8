"""The stream-ordered FP32 vector operations of a PCG iteration
(xp_init_pcg_f32, xp_dot_f32, xp_update_xrz_f32, xp_update_p_f32: the
reference's init_pcg / dot / update_xrz / update_p kernels,
archived/kernels/cg_kernels.cpp:73-158, on device pointers).

Bars (written in each test):
  * every stored vector equals numpy's evaluation of the same formula in
    float64 from the float32 operands, rounded to float32 once -- bit for
    bit (the kernels evaluate in FP64 without fused multiply-add);
  * every reduction (r^T z, r^T r, a^T b) is within n * 2^-52 * sum |a_i b_i|
    of the exactly rounded sum (math.fsum of the exact float32 products);
    per system for a batched matrix;
  * a PCG loop composed on the host from xp_spmv_f32 and these operations
    (the reference's local multi-kernel orchestration,
    archived/host/local_multi_kernel_solver.cpp) follows the oracle's FP64
    golden PCG: the first iterations' alpha, beta, rz, rr within 1e-3
    relative, the iteration count to rr <= tau within 3 %, and the final
    solution within 1e-4 of the golden one.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -52


def f32(rng, n, lo=-1.0, hi=1.0):
    return rng.uniform(lo, hi, n).astype(np.float32)


def sys_of_rows(m, systems):
    """System index of each solve-order row (padding rows: -1)."""
    perm = m.permutation()
    starts = np.cumsum([0] + [s.n for s in systems])
    out = np.full(perm.size, -1)
    real = perm >= 0
    out[real] = np.searchsorted(starts, perm[real], side="right") - 1
    return out


def dot_bound(a, b):
    prod = a.astype(np.float64) * b.astype(np.float64)       # exact
    return math.fsum(prod), a.size * EPS * np.abs(prod).sum() + 1e-300


def check_dot(got, a, b):
    want, tol = dot_bound(a, b)
    assert abs(got - want) <= tol, (got, want, tol)


@pytest.mark.parametrize("batch", [1, 3])
def test_ops_bit_exact_and_reductions(gpu, batch):
    systems = [X.gen_diffusion2d(24 + 8 * k, 20, 1.0, 7 + k) for k in range(batch)]
    m = X.Matrix(systems if batch > 1 else systems[0])
    info = m.info()
    n_pad, nsys = info["n_pad"], info["nsys"]
    assert nsys == batch
    owner = sys_of_rows(m, systems)
    real = owner >= 0
    print(f"batch={batch} n_pad={n_pad} grid={info['grid']}", flush=True)
    rng = np.random.default_rng(42 + batch)
    vec = {k: np.where(real, f32(rng, n_pad), 0).astype(np.float32) for k in ("b", "ax", "x", "ap", "pp")}
    vec["minv"] = np.where(real, f32(rng, n_pad, 0.1, 2.0), 0).astype(np.float32)
    dev = {k: X.DeviceArray.from_numpy(v) for k, v in vec.items()}
    r, z, p = (X.DeviceArray(n_pad, np.float32) for _ in range(3))
    metrics = X.DeviceArray(2 * nsys, np.float64)
    out = X.DeviceArray(nsys, np.float64)
    d64 = np.float64

    # init: r = b - ax, z = minv r, p = z
    m.init_pcg_f32(dev["b"].ptr, dev["ax"].ptr, dev["minv"].ptr, r.ptr, z.ptr, p.ptr, metrics.ptr)
    X.synchronize()
    rw = (vec["b"].astype(d64) - vec["ax"].astype(d64)).astype(np.float32)
    zw = (vec["minv"].astype(d64) * rw.astype(d64)).astype(np.float32)
    np.testing.assert_array_equal(r.numpy(), rw)
    np.testing.assert_array_equal(z.numpy(), zw)
    np.testing.assert_array_equal(p.numpy(), zw)
    got = metrics.numpy()
    for k in range(nsys):
        sel = owner == k
        check_dot(got[2 * k], rw[sel], zw[sel])
        check_dot(got[2 * k + 1], rw[sel], rw[sel])

    # dot
    m.dot_f32(dev["pp"].ptr, dev["ap"].ptr, out.ptr)
    X.synchronize()
    got = out.numpy()
    for k in range(nsys):
        sel = owner == k
        check_dot(got[k], vec["pp"][sel], vec["ap"][sel])

    # update_xrz with p from init: x += alpha p, r -= alpha ap, z = minv r
    alpha = 0.3712345678901234
    m.update_xrz_f32(dev["x"].ptr, p.ptr, r.ptr, dev["ap"].ptr, dev["minv"].ptr, z.ptr, metrics.ptr, alpha)
    X.synchronize()
    xw = (vec["x"].astype(d64) + alpha * zw.astype(d64)).astype(np.float32)
    r2 = (rw.astype(d64) - alpha * vec["ap"].astype(d64)).astype(np.float32)
    z2 = (vec["minv"].astype(d64) * r2.astype(d64)).astype(np.float32)
    np.testing.assert_array_equal(dev["x"].numpy(), xw)
    np.testing.assert_array_equal(r.numpy(), r2)
    np.testing.assert_array_equal(z.numpy(), z2)
    got = metrics.numpy()
    for k in range(nsys):
        sel = owner == k
        check_dot(got[2 * k], r2[sel], z2[sel])
        check_dot(got[2 * k + 1], r2[sel], r2[sel])

    # update_p: p = z + beta p
    beta = -1.2345678912345
    m.update_p_f32(z.ptr, p.ptr, beta)
    X.synchronize()
    pw = (z2.astype(d64) + beta * zw.astype(d64)).astype(np.float32)
    np.testing.assert_array_equal(p.numpy(), pw)


def hip_runtime():
    """The HIP runtime the library is linked against (already loaded)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


def test_ops_on_a_caller_stream(gpu):
    """The operations are ordered on the stream the caller passes."""
    import ctypes
    a = X.gen_diffusion2d(40, 40, 1.0, 3)
    m = X.Matrix(a)
    n_pad = m.info()["n_pad"]
    owner = sys_of_rows(m, [a])
    rng = np.random.default_rng(5)
    va = np.where(owner >= 0, f32(rng, n_pad), 0).astype(np.float32)
    vb = np.where(owner >= 0, f32(rng, n_pad), 0).astype(np.float32)
    da, db = X.DeviceArray.from_numpy(va), X.DeviceArray.from_numpy(vb)
    out = X.DeviceArray(1, np.float64)
    hip = hip_runtime()
    s = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(s)) == 0
    try:
        for _ in range(3):
            m.dot_f32(da.ptr, db.ptr, out.ptr, s.value)
            m.update_p_f32(da.ptr, db.ptr, 0.5, s.value)
        assert hip.hipStreamSynchronize(s) == 0
    finally:
        hip.hipStreamDestroy(s)
    want = [vb]
    for _ in range(3):
        want.append((va.astype(np.float64) + 0.5 * want[-1].astype(np.float64)).astype(np.float32))
    np.testing.assert_array_equal(db.numpy(), want[3])
    # the last dot saw the second update's p
    check_dot(float(out.numpy()[0]), va, want[2])


def composed_pcg(m, a, tau, max_iters):
    """The reference's multi-kernel orchestration on device vectors: the
    scalars alpha = rz / p^T A p and beta = rz_new / rz on the host."""
    info = m.info()
    n_pad = info["n_pad"]
    perm = m.permutation()
    real = perm >= 0
    diag = np.zeros(a.n)
    for r in range(a.n):
        for o in range(a.row_ptr[r], a.row_ptr[r + 1]):
            if a.col_idx[o] == r:
                diag[r] += a.values[o]
    minv_s = np.zeros(n_pad, np.float32)
    minv_s[real] = (1.0 / diag[perm[real]]).astype(np.float32)
    b_s = np.zeros(n_pad, np.float32)
    b_s[real] = a.b[perm[real]].astype(np.float32)
    x_s = np.zeros(n_pad, np.float32)
    x_s[real] = a.x0[perm[real]].astype(np.float32)
    x, b, minv = (X.DeviceArray.from_numpy(v) for v in (x_s, b_s, minv_s))
    r, z, p, ap = (X.DeviceArray(n_pad, np.float32) for _ in range(4))
    metrics = X.DeviceArray(2, np.float64)
    pap = X.DeviceArray(1, np.float64)
    m.spmv_f32(x.ptr, ap.ptr)
    m.init_pcg_f32(b.ptr, ap.ptr, minv.ptr, r.ptr, z.ptr, p.ptr, metrics.ptr)
    X.synchronize()
    rz, rr = metrics.numpy()
    trace = []
    it = 0
    while it < max_iters and rr > tau:
        m.spmv_f32(p.ptr, ap.ptr)
        m.dot_f32(p.ptr, ap.ptr, pap.ptr)
        X.synchronize()
        alpha = rz / pap.numpy()[0]
        m.update_xrz_f32(x.ptr, p.ptr, r.ptr, ap.ptr, minv.ptr, z.ptr, metrics.ptr, alpha)
        X.synchronize()
        rz_new, rr = metrics.numpy()
        beta = rz_new / rz
        m.update_p_f32(z.ptr, p.ptr, beta)
        trace.append((alpha, beta, rz_new, rr))
        rz = rz_new
        it += 1
    X.synchronize()
    xs = x.numpy()
    xn = np.zeros(a.n)
    xn[perm[real]] = xs[real]
    return xn, it, np.array(trace)


def test_composed_pcg_follows_golden(gpu):
    a = X.gen_diffusion2d(32, 32, 1.0, 20261204)
    m = X.Matrix(a)
    g0 = O.golden(a, tau=1e-30, max_iters=1)
    tau = 1e-10 * g0["trace"][0, 4]           # rr reduced by 1e10 (well inside FP32 reach)
    g = O.golden(a, tau=tau, max_iters=5000)
    assert g["converged"]
    x, it, tr = composed_pcg(m, a, tau, 5000)
    print(f"composed FP32 loop: {it} iterations, golden {g['iterations']}", flush=True)
    k = min(10, len(tr))
    np.testing.assert_allclose(tr[:k], g["trace"][:k, 1:5], rtol=1e-3)
    assert abs(it - g["iterations"]) <= max(2, 0.03 * g["iterations"])
    assert np.abs(x - g["x"]).max() <= 1e-4 * max(1.0, np.abs(g["x"]).max())
