"""CPU oracle (pure Python) of the strength-aggregated 8-row blocks.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the product).  The
reference has no block preconditioner; this is the independent restatement
of the library's greedy strength aggregation that its C++ builder
(xp_strength_blocking) is checked against:

  s_ij = |a_ij| / sqrt(a_ii a_jj)   (1 instead of 1/sqrt(a_ii) where a_ii <= 0
                                     or missing; the first diagonal entry)
  pass 1: seeds in row order; a block grows by the unassigned neighbour with
          the largest summed strength to the block's rows (ties: the smaller
          row) until it has 8 rows or no unassigned neighbour;
  pass 2: the rows of blocks smaller than 8 are pooled and regrouped the
          same way among themselves;
  pack:   blocks still smaller than 8 are merged first-fit, in creation
          order, into blocks of at most 8 rows;
  rows ascending inside a block, blocks ordered by their first row.

rule 2 (strength-ordered): the same passes with seeds in decreasing order of
the row's total off-diagonal strength t_i = sum_j s_ij (CSR order; ties: the
smaller row) and candidates scored by (summed strength to the block) / t_j;
the remaining partial blocks are packed in order of their smallest row, each
into the most recent of the last 16 packed blocks with room (else a new one).

Pinning (tests/test_blocking.py): partition and size invariants; on a 1-D
chain (tridiagonal) the blocks are the consecutive runs of 8; the greedy
choice is checked step by step against brute force on small matrices.
"""
from __future__ import annotations

import math


def strengths(row_ptr, col_idx, values):
    n = len(row_ptr) - 1
    dis = [1.0] * n
    for r in range(n):
        for o in range(row_ptr[r], row_ptr[r + 1]):
            if col_idx[o] == r:
                if values[o] > 0:
                    dis[r] = 1.0 / math.sqrt(values[o])
                break
    return [abs(values[o]) * dis[r] * dis[col_idx[o]]
            for r in range(n) for o in range(row_ptr[r], row_ptr[r + 1])]


def _grow(seeds, row_ptr, col_idx, s, total, eligible, owner, out):
    for seed in seeds:
        if not eligible[seed] or owner[seed] >= 0:
            continue
        bid = len(out)
        blk = [seed]
        owner[seed] = bid
        cand = []          # [row, summed strength], in first-seen order

        def add(v):
            for o in range(row_ptr[v], row_ptr[v + 1]):
                j = col_idx[o]
                if j == v or not eligible[j] or owner[j] >= 0:
                    continue
                for c in cand:
                    if c[0] == j:
                        c[1] += s[o]
                        break
                else:
                    cand.append([j, s[o]])

        add(seed)
        while len(blk) < 8:
            best, bw, bi = -1, -1.0, -1
            for i, (j, w) in enumerate(cand):
                if owner[j] >= 0:
                    continue
                if total is not None:
                    w = w / total[j] if total[j] > 0 else 0.0
                if w > bw or (w == bw and j < best):
                    best, bw, bi = j, w, i
            if best < 0:
                break
            del cand[bi]
            owner[best] = bid
            blk.append(best)
            add(best)
        out.append(blk)


def strength_blocking(row_ptr, col_idx, values, rule=1):
    """Blocks as a list of row lists (ascending rows, ordered by first row).
    rule 2: seeds by decreasing total off-diagonal strength t_i (summed in
    CSR order; ties: the smaller row), candidates scored by (summed strength
    to the block) / t_j."""
    n = len(row_ptr) - 1
    if n == 0:
        return []
    s = strengths(row_ptr, col_idx, values)
    seeds, total = list(range(n)), None
    if rule == 2:
        total = [0.0] * n
        for r in range(n):
            for o in range(row_ptr[r], row_ptr[r + 1]):
                if col_idx[o] != r:
                    total[r] += s[o]
        seeds.sort(key=lambda r: -total[r])          # stable: ties keep row order
    owner = [-1] * n
    first = []
    _grow(seeds, row_ptr, col_idx, s, total, [True] * n, owner, first)
    full, pool = [], [False] * n
    for blk in first:
        if len(blk) == 8:
            full.append(blk)
        else:
            for r in blk:
                pool[r] = True
                owner[r] = -1
    second = []
    _grow(seeds, row_ptr, col_idx, s, total, pool, owner, second)
    partial = []
    for blk in second:
        (full if len(blk) == 8 else partial).append(blk)
    packed = []
    if rule == 1:
        for blk in partial:
            for p in packed:
                if len(p) + len(blk) <= 8:
                    p.extend(blk)
                    break
            else:
                packed.append(list(blk))
    else:
        # pieces by smallest row; each into the most recent of the last 16
        # packed blocks with room, else a new block
        partial = sorted((sorted(b) for b in partial), key=lambda b: b[0])
        for blk in partial:
            for p in reversed(packed[-16:]):
                if len(p) + len(blk) <= 8:
                    p.extend(blk)
                    break
            else:
                packed.append(list(blk))
    full.extend(packed)
    full = [sorted(b) for b in full]
    full.sort(key=lambda b: b[0])
    return full
