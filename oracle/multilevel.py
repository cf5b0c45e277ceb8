"""CPU oracle (FP64 numpy) of the multilevel aggregation preconditioner.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the product).  The
reference has no multilevel preconditioner: its Jacobi-PCG loop takes the
inverse diagonal (src/CgSolverGolden.hpp:33-40), which the block-Jacobi
preconditioner generalises to 8x8 blocks and this one extends with coarse
levels.  This module is the independent CPU restatement the GPU hierarchy
and V-cycle are checked against:

  strong_graph   -- the strength graph: j != i is a strong neighbour of i
                    when a_ij^2 >= theta^2 |a_ii| |a_jj| (FP64, squared so
                    that no square root enters)
  priority_key   -- a row's priority in the root selection: the high 32
                    bits of MurmurHash3's 64-bit finaliser (fmix64) of
                    i + salt * 0x9E3779B97F4A7C15, above i + 1 (unique keys)
  plain_aggregation
                 -- aggregates of a level: roots are a maximal distance-2
                    independent set of the strength graph, chosen in Luby
                    rounds (an undecided row becomes a root when its key is
                    the largest among the undecided rows within strong
                    distance 2; rows within distance 2 of a root are then
                    excluded); root k (index order) is aggregate k; a
                    root's strong neighbours join it; every other row joins
                    the aggregate of its strongest joined neighbour (largest
                    |a_ij|, then the smallest j); rows left over become
                    aggregates of their own, in index order
  members        -- each aggregate's rows, ascending
  pieces         -- 8-row smoother blocks cut from aggregates: ceil(size/8)
                    pieces of consecutive rows per aggregate, sizes as
                    equal as possible (the first size % pieces one longer)
  galerkin       -- the coarse matrix P^T A P (P piecewise constant on the
                    aggregates): row u = aggregate u, entry (u, v) = sum of
                    a_ij over rows i of u and columns j of v, equal columns
                    summed in the order u's rows (ascending) and their CSR
                    entries give them, columns ascending
  exact_blocks   -- the exact inverses of a matrix's diagonal blocks on its
                    smoother blocks (padding slots identity), from the
                    matrix's FP32-rounded values, as the device forms its
                    coarse smoothers
  Hierarchy      -- levels (matrix, smoother blocks, block operators,
                    aggregates) down to a dense coarsest matrix
  vcycle         -- one symmetric V-cycle z = V r: z = omega M r;
                    z += gamma P V(P^T (r - A z)); z += omega M (r - A z);
                    the coarsest level solved exactly
  two_level_dense-- the same cycle for two levels written as one dense
                    matrix, V = (I - B A)(B + C (I - A B)) + B with
                    B = omega M and C = gamma P A_c^-1 P^T
  pcg            -- FP64 PCG with z = V(r), stopped on r^T r <= tau

Pinning (tests/test_multilevel_oracle.py): plain_aggregation against its
defining properties checked by breadth-first search (roots pairwise at
strong distance >= 3, every row within distance 2 of a root, the join
rules) and against a brute-force restatement of the rounds on small graphs
(all-pairs distances instead of max propagation); the properties hold for
any injective key, and priority_key's keys are checked to be distinct (the
device computes the same function: its aggregates equal these bit for bit);
galerkin against scipy's P^T A P (an
independent formula, to FP64 rounding) and bit for bit against a
hand-computed case; vcycle against two_level_dense; the cycle's symmetry and
positive definiteness (a preconditioner PCG needs), checked on its dense
matrix; pcg against numpy's solve on small systems.
"""
from __future__ import annotations

import numpy as np

from block_mlp import _dot, stored_blocks


MASK64 = (1 << 64) - 1
GOLDEN64 = 0x9E3779B97F4A7C15


def fmix64(x):
    """MurmurHash3's 64-bit finaliser (public domain, A. Appleby)."""
    x &= MASK64
    x ^= x >> 33
    x = (x * 0xFF51AFD7ED558CCD) & MASK64
    x ^= x >> 33
    x = (x * 0xC4CEB9FE1A85EC53) & MASK64
    x ^= x >> 33
    return x


def priority_key(i, salt):
    """(high 32 bits of fmix64(i + salt * golden)) << 32 | (i + 1)."""
    return ((fmix64(int(i) + int(salt) * GOLDEN64) >> 32) << 32) | (int(i) + 1)


def strong_graph(row_ptr, col_idx, values, theta):
    """Strong neighbours of every row: (indptr, indices, |a_ij|) in CSR
    order, j != i, a_ij^2 >= theta^2 |a_ii| |a_jj| with a_ii the sum of the
    row's diagonal entries (FP64)."""
    row_ptr = np.asarray(row_ptr, np.int64)
    col_idx = np.asarray(col_idx, np.int64)
    values = np.asarray(values, np.float64)
    n = len(row_ptr) - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    diag = np.zeros(n)
    on = col_idx == rows
    for e in np.flatnonzero(on):  # CSR order
        diag[rows[e]] += values[e]
    d = np.abs(diag)
    th2 = float(theta) * float(theta)
    keep = (~on) & (values * values >= th2 * (d[rows] * d[col_idx]))
    cnt = np.bincount(rows[keep], minlength=n)
    indptr = np.concatenate([[0], np.cumsum(cnt)])
    return indptr, col_idx[keep], np.abs(values[keep])


def _nbmax(indptr, indices, v):
    """max of v over {i} and i's strong neighbours."""
    out = v.copy()
    has = np.diff(indptr) > 0
    if indices.size:
        red = np.maximum.reduceat(v[indices], indptr[:-1][has])
        out[has] = np.maximum(out[has], red)
    return out


def plain_aggregation(row_ptr, col_idx, values, theta, salt, with_roots=False):
    """(agg, nagg, rounds): the aggregate of every row, the aggregate count
    and the number of root-selection rounds (module docstring); with_roots:
    the roots (ascending) as a fourth item."""
    sp_, si, sv = strong_graph(row_ptr, col_idx, values, theta)
    n = len(row_ptr) - 1
    keys = np.array([priority_key(i, salt) for i in range(n)], dtype=np.uint64)
    state = np.zeros(n, np.int8)  # 0 undecided, 1 root, 2 excluded
    rounds = 0
    while (state == 0).any():
        rounds += 1
        k = np.where(state == 0, keys, np.uint64(0))
        m2 = _nbmax(sp_, si, _nbmax(sp_, si, k))
        state[(state == 0) & (m2 == keys)] = 1
        near = _nbmax(sp_, si, _nbmax(sp_, si, (state == 1).astype(np.int8))) > 0
        state[(state == 0) & near] = 2
    roots = np.flatnonzero(state == 1)
    agg1 = np.full(n, -1, np.int64)
    agg1[roots] = np.arange(roots.size)
    for i in range(n):
        if state[i] == 1:
            continue
        for e in range(sp_[i], sp_[i + 1]):
            if state[si[e]] == 1:
                agg1[i] = agg1[si[e]]
    agg = agg1.copy()
    for i in np.flatnonzero(agg1 < 0):
        best, bj = -1.0, -1
        for e in range(sp_[i], sp_[i + 1]):
            j = si[e]
            if agg1[j] >= 0 and (sv[e] > best or (sv[e] == best and j < bj)):
                best, bj = sv[e], j
        if bj >= 0:
            agg[i] = agg1[bj]
    left = np.flatnonzero(agg < 0)
    agg[left] = roots.size + np.arange(left.size)
    out = (agg.astype(np.int32), int(roots.size + left.size), rounds)
    return out + (roots,) if with_roots else out


def members(agg, nagg):
    """Each aggregate's rows, ascending: (ptr, rows)."""
    agg = np.asarray(agg, np.int64)
    order = np.argsort(agg, kind="stable")
    ptr = np.concatenate([[0], np.cumsum(np.bincount(agg, minlength=nagg))])
    return ptr.astype(np.int32), order.astype(np.int32)


def pieces(agg, nagg):
    """8-row smoother blocks cut from aggregates: aggregate a (rows
    ascending) gives ceil(size / 8) pieces of consecutive rows, the first
    size % pieces of them one row longer; pieces in aggregate order, as
    units x 8 slots with -1 padding."""
    mp, mr = members(agg, nagg)
    out = []
    for a in range(nagg):
        rows = mr[mp[a]:mp[a + 1]]
        k = -(-len(rows) // 8)
        q, rem = divmod(len(rows), k)
        o = 0
        for p in range(k):
            ln = q + (1 if p < rem else 0)
            out.append(list(rows[o:o + ln]) + [-1] * (8 - ln))
            o += ln
    return np.array(out, np.int32).reshape(-1, 8)


def galerkin(row_ptr, col_idx, values, agg, nagg):
    """P^T A P of a CSR over the aggregates agg (row -> aggregate): returns
    (row_ptr, col_idx, values) of the coarse CSR (natural order = aggregate
    ids).  Within coarse row u, the entries of u's rows are taken row by row
    (ascending) and in CSR order; the first entry of a coarse column starts
    its sum and the later ones are added in that order."""
    row_ptr = np.asarray(row_ptr)
    col_idx = np.asarray(col_idx)
    values = np.asarray(values, dtype=np.float64)
    agg = np.asarray(agg, np.int64)
    mp, mr = members(agg, nagg)
    rp, ci, va = [0], [], []
    for u in range(nagg):
        acc = {}
        for r in mr[mp[u]:mp[u + 1]]:
            for e in range(row_ptr[r], row_ptr[r + 1]):
                c = int(agg[col_idx[e]])
                acc[c] = acc[c] + values[e] if c in acc else values[e]
        for c in sorted(acc):
            ci.append(c)
            va.append(acc[c])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float64)


def slots_agg(slots):
    """The aggregates given by 8-row blocks: row -> its block."""
    slots = np.asarray(slots).reshape(-1, 8)
    n = int(slots.max()) + 1
    agg = np.full(n, -1, np.int64)
    for u, grp in enumerate(slots):
        for r in grp:
            if r >= 0:
                agg[r] = u
    return agg.astype(np.int32), len(slots)


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
    """A level: matrix A, smoother blocks (8-row slots) with their block
    operators, and the aggregates agg (row -> row of the next level) that
    define its piecewise-constant prolongation P."""

    def __init__(self, A, slots, blocks, agg=None, nagg=None):
        import scipy.sparse as sp
        self.A = A.tocsr()
        self.slots = np.asarray(slots).reshape(-1, 8)
        n = self.A.shape[0]
        self.idx = np.full(self.slots.shape, n)
        for k, grp in enumerate(self.slots):
            g = [int(r) for r in grp if r >= 0]
            self.idx[k, :len(g)] = g
        self.blocks = np.asarray(blocks, dtype=np.float64)
        if agg is None:
            agg, nagg = slots_agg(self.slots)
        self.agg = np.asarray(agg, np.int64)
        self.P = sp.csr_matrix((np.ones(n), (np.arange(n), self.agg)), shape=(n, int(nagg)))

    def smooth(self, r):
        """M r with the level's block operators."""
        n = r.size
        rp = np.append(r, 0.0)[self.idx]
        z = np.zeros(n + 1)
        z[self.idx] = np.einsum("kij,kj->ki", self.blocks, rp)
        return z[:n]


class Hierarchy:
    """levels[l]: Level (A_l, its smoother blocks and operators, its
    aggregates); coarse: the dense coarsest matrix (the Galerkin product
    over the last level's aggregates)."""

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
