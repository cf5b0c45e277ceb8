"""CPU oracle (FP64 numpy) of the multilevel aggregation preconditioner.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the product).  The
reference has no multilevel preconditioner: its Jacobi-PCG loop takes the
inverse diagonal (src/CgSolverGolden.hpp:33-40), which the block-Jacobi
preconditioner generalises to 8x8 blocks and this one extends with coarse
levels.  This module is the independent CPU restatement the GPU hierarchy
and V-cycle are checked against:

  galerkin       -- the coarse matrix P^T A P over 8-row aggregates (P
                    piecewise constant): row u = aggregate u, entry (u, v) =
                    sum of a_ij over rows i of u and columns j of v, equal
                    columns summed in the order the aggregate's rows (slot
                    order) and their CSR entries give them, columns ascending
  exact_blocks   -- the exact inverses of a matrix's diagonal blocks on its
                    aggregates (padding slots identity), from the matrix's
                    FP32-rounded values, as the device forms its coarse
                    smoothers
  Hierarchy      -- levels (matrix, aggregates, block operators) down to a
                    dense coarsest matrix
  vcycle         -- one symmetric V-cycle z = V r: z = omega M r;
                    z += gamma P V(P^T (r - A z)); z += omega M (r - A z);
                    the coarsest level solved exactly
  two_level_dense-- the same cycle for two levels written as one dense
                    matrix, V = (I - B A)(B + C (I - A B)) + B with
                    B = omega M and C = gamma P A_c^-1 P^T
  pcg            -- FP64 PCG with z = V(r), stopped on r^T r <= tau

Pinning (tests/test_multilevel_oracle.py): galerkin against scipy's P^T A P
(an independent formula, to FP64 rounding) and bit for bit against a
hand-computed case; vcycle against two_level_dense; the cycle's symmetry
and positive definiteness (a preconditioner PCG needs), checked on its
dense matrix; pcg against numpy's solve on small systems.
"""
from __future__ import annotations

import numpy as np

from block_mlp import _dot, stored_blocks


def galerkin(row_ptr, col_idx, values, slots):
    """P^T A P of a CSR over aggregates slots (units x 8 rows, -1 padding):
    returns (row_ptr, col_idx, values) of the coarse CSR (natural order =
    aggregate ids).  Within coarse row u, the entries of u's rows are taken
    slot by slot and in CSR order; the first entry of a coarse column starts
    its sum and the later ones are added in that order."""
    row_ptr = np.asarray(row_ptr)
    col_idx = np.asarray(col_idx)
    values = np.asarray(values, dtype=np.float64)
    slots = np.asarray(slots).reshape(-1, 8)
    n = len(row_ptr) - 1
    agg = np.full(n, -1, dtype=np.int64)
    for u, grp in enumerate(slots):
        for r in grp:
            if r >= 0:
                agg[r] = u
    rp, ci, va = [0], [], []
    for grp in slots:
        acc = {}
        for r in grp:
            if r < 0:
                continue
            for e in range(row_ptr[r], row_ptr[r + 1]):
                c = int(agg[col_idx[e]])
                acc[c] = acc[c] + values[e] if c in acc else values[e]
        for c in sorted(acc):
            ci.append(c)
            va.append(acc[c])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float64)


def to_dense(row_ptr, col_idx, values, n):
    A = np.zeros((n, n))
    for i in range(n):
        for e in range(row_ptr[i], row_ptr[i + 1]):
            A[i, col_idx[e]] += values[e]
    return A


def exact_blocks(A, slots, f32_values=True):
    """Inverses of the diagonal blocks of A (scipy CSR) on the aggregates,
    8 x 8 each with identity on padding slots; from A's values rounded to
    FP32 when f32_values (the device's coarse smoothers read its FP32
    value stream)."""
    slots = np.asarray(slots).reshape(-1, 8)
    out = np.zeros((len(slots), 8, 8))
    for k, grp in enumerate(slots):
        g = [int(r) for r in grp if r >= 0]
        sub = A[g][:, g].toarray()
        if f32_values:
            sub = sub.astype(np.float32).astype(np.float64)
        out[k] = np.eye(8)
        out[k][:len(g), :len(g)] = np.linalg.inv(sub)
    return out


class Level:
    def __init__(self, A, slots, blocks):
        self.A = A.tocsr()
        self.slots = np.asarray(slots).reshape(-1, 8)
        n = self.A.shape[0]
        self.idx = np.full(self.slots.shape, n)
        for k, grp in enumerate(self.slots):
            g = [int(r) for r in grp if r >= 0]
            self.idx[k, :len(g)] = g
        self.blocks = np.asarray(blocks, dtype=np.float64)
        nb = len(self.slots)
        agg = np.zeros(n, np.int64)
        for k, grp in enumerate(self.slots):
            for r in grp:
                if r >= 0:
                    agg[r] = k
        import scipy.sparse as sp
        self.P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nb))

    def smooth(self, r):
        """M r with the level's block operators."""
        n = r.size
        rp = np.append(r, 0.0)[self.idx]
        z = np.zeros(n + 1)
        z[self.idx] = np.einsum("kij,kj->ki", self.blocks, rp)
        return z[:n]


class Hierarchy:
    """levels[l]: Level (A_l, its aggregates, its block operators); coarse:
    the dense coarsest matrix (the Galerkin product over the last level's
    aggregates)."""

    def __init__(self, levels, coarse, omega, gamma):
        self.levels = levels
        self.coarse = np.asarray(coarse, dtype=np.float64)
        self.coarse_inv = np.linalg.inv(self.coarse)
        self.omega = float(omega)
        self.gamma = float(gamma)


def vcycle(H, r, level=0):
    """z = V r on level `level` (natural order of that level)."""
    if level == len(H.levels):
        return H.coarse_inv @ r
    L = H.levels[level]
    z = H.omega * L.smooth(r)
    e = vcycle(H, L.P.T @ (r - L.A @ z), level + 1)
    t = z + H.gamma * (L.P @ e)
    return t + H.omega * L.smooth(r - L.A @ t)


def two_level_dense(A, M, P, Ac, omega, gamma):
    """The two-level cycle as one matrix: B = omega M, C = gamma P Ac^-1 P^T,
    V = (I - B A)(B + C (I - A B)) + B."""
    n = A.shape[0]
    B = omega * M
    C = gamma * (P @ np.linalg.solve(Ac, P.T))
    I = np.eye(n)
    return (I - B @ A) @ (B + C @ (I - A @ B)) + B


def pcg(A, b, prec, tau, max_iters=100000):
    """FP64 PCG, z = prec(r), stopped when r^T r <= tau: (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz, rr = _dot(r, z), _dot(r, r)
    it = 0
    while it < max_iters and rr > tau:
        ap = A @ p
        alpha = rz / _dot(p, ap)
        x += alpha * p
        r -= alpha * ap
        z = prec(r)
        rzn, rr = _dot(r, z), _dot(r, r)
        it += 1
        if rr <= tau:
            break
        p = z + (rzn / rz) * p
        rz = rzn
    return x, it


def stored(blocks):
    """The operators the device applies for full blocks (FP16 factors)."""
    return stored_blocks(blocks)
