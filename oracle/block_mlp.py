"""CPU oracle (FP64 numpy) of the learned 8x8 block-Jacobi preconditioner.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks, never by
the product).  The reference has no learned preconditioner; this module is
the independent CPU restatement its GPU kernels are checked against:

  features_from_csr  -- block features of S = D^-1/2 A D^-1/2, Ahat = S_bb,
                        H = Ahat^2 + LAMBDA G, D^-1/2 per row, for natural
                        8-row groups or any blocks of at most 8 rows
  forward            -- the MLP 48 -> 64 -> 64 -> 48 (ReLU, ReLU, linear)
  l0_from_feat       -- L0 = chol(Ahat^-1), the exact block inverse's lower
                        factor (numpy Cholesky of numpy's inverse); the
                        network's output is a correction: L = L0 + tril(out)
  loss_grad          -- mean block-SPAI loss and its analytic gradient
  blocks_from_out    -- M_b = D^-1/2 (L L^T + 1e-3 I) D^-1/2
  stored_blocks      -- the operator the GPU applies for blocks M_b: s = sqrt(diag M_b),
                        C = chol(S^-1 M_b S^-1) rounded to FP16, S C C^T S
  block_pcg          -- FP64 PCG with a block-diagonal preconditioner
  block_pcg_accurate -- the same PCG in the library's accurate arithmetic
                        (r, z rounded to FP32; p, A p, x FP64; A p by the
                        recurrence A z + beta A p) stopped on the TRUE
                        residual: when the recursive r^T r meets tau but
                        ||b - A x||^2 does not, r and z are replaced by the
                        true residual and the recurrence restarts (p = z)

Pinning: loss_grad against central finite differences of loss (FP64,
tests/test_block_oracle.py); block_pcg against the golden Jacobi-PCG when
the blocks are diagonal, and against finite termination with exact block
inverses on block-diagonal matrices; block_pcg_accurate without its FP32
rounding against block_pcg (same iterates up to FP64 rounding), and with it
against the properties the rounding must keep (r and z FP32-representable,
the stop on the true residual); features_from_csr against dense closed
forms on hand-built blocks.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_H, N_OUT = 48, 64, 48
SHAPES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H, N_OUT)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
EPS_M = 1e-3
LAMBDA = 0.5
TRIL = [(r, c) for r in range(8) for c in range(r + 1)]


def _dot(a, b):
    """a . b in an order that does not depend on the CPU's vector width or
    BLAS: the products, then column sums of a 1024-wide reshape added row
    after row (elementwise adds), then the 1024 partial sums added exactly
    (math.fsum).  The PCG iteration counts below are sensitive to last-bit
    differences in these sums; this keeps them the same on every machine."""
    v = np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)
    pad = (-v.shape[0]) % 1024
    if pad:
        v = np.concatenate([v, np.zeros(pad)])
    return math.fsum(v.reshape(-1, 1024).sum(axis=0).tolist())


def _block_apply(blocks, rp):
    """z_bi = sum_j M_bij rp_bj, the eight terms added in order j = 0..7
    (elementwise operations only, so no CPU-dependent summation order)."""
    z = blocks[:, :, 0] * rp[:, None, 0]
    for j in range(1, 8):
        z = z + blocks[:, :, j] * rp[:, None, j]
    return z


def unpack(flat):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def natural_groups(n):
    return [list(range(r, min(r + 8, n))) for r in range(0, n, 8)]


def features_from_csr(row_ptr, col_idx, values, groups=None):
    """Features of the blocks `groups` (lists of at most 8 rows; slot j of a
    block is its j-th row, padding slots after them), natural 8-row groups
    when None.  Returns feat (nb x 48), ahat, hmat (nb x 8 x 8), scale (nb x 8)."""
    n = len(row_ptr) - 1
    if groups is None:
        groups = natural_groups(n)
    nb = len(groups)
    dis = np.ones(n)
    for r in range(n):
        for o in range(row_ptr[r], row_ptr[r + 1]):
            if col_idx[o] == r:
                if values[o] > 0:
                    dis[r] = 1.0 / np.sqrt(values[o])
                break
    where = {}
    for b, g in enumerate(groups):
        for j, r in enumerate(g):
            where[r] = (b, j)
    feat = np.zeros((nb, N_IN))
    ahat = np.zeros((nb, 8, 8))
    hmat = np.zeros((nb, 8, 8))
    scale = np.ones((nb, 8))
    for b, g in enumerate(groups):
        m = len(g)
        a = np.zeros((8, 8))
        offabs = np.zeros(8)
        off = [dict() for _ in range(8)]
        for i, r in enumerate(g):
            for o in range(row_ptr[r], row_ptr[r + 1]):
                c = col_idx[o]
                v = values[o] * dis[r] * dis[c]
                bc, jc = where[c]
                if bc == b:
                    a[i, jc] += v
                else:
                    offabs[i] += abs(v)
                    off[i][c] = off[i].get(c, 0.0) + v
        for i in range(m, 8):
            a[i, i] = 1.0
        scale[b, :m] = dis[list(g)]
        gm = np.zeros((8, 8))
        for i in range(8):
            for j in range(8):
                gm[i, j] = sum(v * off[j].get(k, 0.0) for k, v in off[i].items())
        ahat[b] = a
        hmat[b] = a @ a + LAMBDA * gm
        feat[b, :36] = a[np.triu_indices(8)]
        feat[b, 36:44] = offabs
    return feat, ahat, hmat, scale


def bf16(x):
    """Round to BF16 (nearest, ties to even), as the GPU kernels do before
    their matrix products."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(x):
    return x


def forward(params, feat, keep=False, round_bf16=False):
    """round_bf16: round the inputs of every matrix product (activations and
    weights) to BF16, where the device's MFMA GEMMs round them."""
    p = unpack(params)
    r = bf16 if round_bf16 else _ident
    z1 = r(feat) @ r(p["W1"]) + p["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = r(h1) @ r(p["W2"]) + p["b2"]
    h2 = np.maximum(z2, 0.0)
    out = r(h2) @ r(p["W3"]) + p["b3"]
    if keep:
        return out, (feat, z1, h1, z2, h2)
    return out


def l_from_out(out):
    n = out.shape[0]
    L = np.zeros((n, 8, 8))
    for t, (r, c) in enumerate(TRIL):
        L[:, r, c] = out[:, t]
    return L


def ahat_from_feat(feat):
    A = np.zeros((feat.shape[0], 8, 8))
    iu = np.triu_indices(8)
    A[:, iu[0], iu[1]] = feat[:, :36]
    A[:, iu[1], iu[0]] = feat[:, :36]
    return A


def l0_from_feat(feat):
    """Lower Cholesky factor of Ahat^-1 per block (identity where Ahat is not
    positive definite), Ahat rebuilt from the feature row."""
    A = ahat_from_feat(np.asarray(feat, dtype=np.float64))
    L0 = np.zeros_like(A)
    for k in range(A.shape[0]):
        try:
            np.linalg.cholesky(A[k])
            L0[k] = np.linalg.cholesky(np.linalg.inv(A[k]))
        except np.linalg.LinAlgError:
            L0[k] = np.eye(8)
    return L0


def factor(out, feat):
    return l0_from_feat(feat) + l_from_out(out)


def loss_grad(params, feat, ahat, hmat, round_bf16=False):
    """Mean over samples of tr(M H M^T) - 2 tr(M A) + 8, M = L L^T + eps I.
    round_bf16: the device's rounding points -- BF16 inputs of the forward
    products and of the two input-gradient products (output gradient,
    hidden gradient, weights); weight gradients stay FP64 here (FP32 on the
    device)."""
    rnd = bf16 if round_bf16 else _ident
    out, (x, z1, h1, z2, h2) = forward(params, feat, keep=True, round_bf16=round_bf16)
    n = out.shape[0]
    L = factor(out, feat)
    M = L @ np.transpose(L, (0, 2, 1)) + EPS_M * np.eye(8)
    MH = M @ hmat
    loss = np.einsum("nij,nij->n", MH, M) - 2.0 * np.einsum("nij,nji->n", M, ahat) + 8.0
    D = 2.0 * (MH - ahat) / n
    dL = (D + np.transpose(D, (0, 2, 1))) @ L
    dout = np.zeros_like(out)
    for t, (r, c) in enumerate(TRIL):
        dout[:, t] = dL[:, r, c]
    p = unpack(params)
    g = {}
    g["W3"] = h2.T @ dout
    g["b3"] = dout.sum(0)
    dh2 = (rnd(dout) @ rnd(p["W3"]).T) * (z2 > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (rnd(dh2) @ rnd(p["W2"]).T) * (z1 > 0)
    g["W1"] = x.T @ dh1
    g["b1"] = dh1.sum(0)
    return float(loss.mean()), pack(g)


def blocks_from_out(out, scale, feat):
    L = factor(out, feat)
    M = L @ np.transpose(L, (0, 2, 1)) + EPS_M * np.eye(8)
    return M * scale[:, :, None] * scale[:, None, :]


def stored_blocks(blocks):
    """What the device stores and applies for SPD blocks (nb x 8 x 8):
    S C C^T S with s_i = sqrt(M_ii) and C the Cholesky factor of the
    unit-diagonal S^-1 M S^-1 rounded to FP16 (a block that is not positive
    definite keeps its diagonal)."""
    blocks = np.asarray(blocks, dtype=np.float64)
    out = np.empty_like(blocks)
    for k, M in enumerate(blocks):
        d = np.diag(M)
        s = np.where(d > 0, np.sqrt(np.where(d > 0, d, 1.0)), 1.0)
        N = M / np.outer(s, s)
        try:
            if not np.all(d > 0):
                raise np.linalg.LinAlgError
            C = np.linalg.cholesky(N).astype(np.float16).astype(np.float64)
        except np.linalg.LinAlgError:
            C = np.eye(8)
        out[k] = np.outer(s, s) * (C @ C.T)
    return out


def block_pcg(A, b, blocks, tau, max_iters=100000, groups=None):
    """FP64 PCG with z_b = M_b r_b on the blocks `groups` (natural 8-row
    groups when None; blocks: nb x 8 x 8 in slot order).
    Returns (x, iterations, rr trace)."""
    n = A.shape[0]
    if groups is None:
        groups = natural_groups(n)
    nb = len(groups)
    idx = np.full((nb, 8), n)                       # padding slots read / write a dummy entry
    for k, g in enumerate(groups):
        idx[k, :len(g)] = g

    def apply(r):
        rp = np.append(r, 0.0)[idx]
        z = np.zeros(n + 1)
        z[idx] = _block_apply(blocks, rp)
        return z[:n]

    x = np.zeros(n)
    r = b.copy()
    z = apply(r)
    p = z.copy()
    rz, rr = _dot(r, z), _dot(r, r)
    trace = []
    it = 0
    while it < max_iters and rr > tau:
        ap = A @ p
        alpha = rz / _dot(p, ap)
        x += alpha * p
        r -= alpha * ap
        z = apply(r)
        rzn, rr = _dot(r, z), _dot(r, r)
        trace.append(rr)
        it += 1
        if rr <= tau:
            break
        p = z + (rzn / rz) * p
        rz = rzn
    return x, it, np.array(trace)


def block_pcg_accurate(A, b, blocks, tau, max_iters=200000, groups=None, round32=True):
    """PCG in the accurate iteration's arithmetic with the operator blocks:
    r and z stored in FP32 (each rounded once from FP64 arithmetic; z from
    the FP64 block product), p, A p and x in FP64, A p carried by the
    recurrence A p = A z + beta A p_old, dots in FP64.  The recursion stops
    when r^T r <= tau; the true residual b - A x is then checked, and when
    its square is above tau, r = fp32(b - A x), z = M r and the recurrence
    restarts (p = z).  round32=False keeps every vector in FP64.
    Returns (x, iterations, restarts)."""
    n = A.shape[0]
    if groups is None:
        groups = natural_groups(n)
    nb = len(groups)
    idx = np.full((nb, 8), n)
    for k, g in enumerate(groups):
        idx[k, :len(g)] = g
    blocks = np.asarray(blocks, dtype=np.float64)

    def apply(r):
        rp = np.append(r, 0.0)[idx]
        z = np.zeros(n + 1)
        z[idx] = _block_apply(blocks, rp)
        return z[:n]

    f = (lambda v: v.astype(np.float32).astype(np.float64)) if round32 else (lambda v: v)
    x = np.zeros(n)
    it = restarts = 0
    r = f(b - A @ x)
    while True:
        z = f(apply(r))
        p = z.copy()
        ap = A @ z
        rz, rr = _dot(r, z), _dot(r, r)
        while rr > tau and it < max_iters:
            alpha = rz / _dot(p, ap)
            x += alpha * p
            r = f(r - alpha * ap)
            z = f(apply(r))
            rzn, rr = _dot(r, z), _dot(r, r)
            it += 1
            if rr <= tau:
                break
            beta = rzn / rz
            rz = rzn
            p = z + beta * p
            ap = A @ z + beta * ap
        rt = b - A @ x
        if _dot(rt, rt) <= tau or it >= max_iters:
            return x, it, restarts
        r = f(rt)
        restarts += 1
