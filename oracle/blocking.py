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

rule 3 (parallel strength-ordered, the device aggregation): the rule-2
ingredients (t_i, the fraction score) in rounds that every row can take at
once -- see parallel_blocking.

Pinning (tests/test_blocking.py): partition and size invariants; on a 1-D
chain (tridiagonal) the blocks are the consecutive runs of 8; the greedy
choice is checked step by step against brute force on small matrices; for
rule 3, the seeds of each round are checked against a brute-force
distance-2 maximum search and every growth step against the choice rule
replayed from scratch.
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


# ---------------------------------------------------------------- rule 3
GROW_STEPS = 8      # growth steps per round
PACK_WINDOW = 16    # packing looks at the last 16 packed blocks of its segment
PACK_SEGMENT = 32   # pieces are packed in independent runs of 32 (one device thread each)


def _f32_bits(v):
    """Bits of v rounded to the nearest FP32 (v >= 0): order-preserving."""
    import struct
    return struct.unpack("<I", struct.pack("<f", v))[0]


def totals(row_ptr, col_idx, s):
    n = len(row_ptr) - 1
    t = [0.0] * n
    for r in range(n):
        for o in range(row_ptr[r], row_ptr[r + 1]):
            if col_idx[o] != r:
                t[r] += s[o]
    return t


def _better(a, b, t):
    """Row a precedes row b in the seed order (larger t, then smaller row)."""
    return b < 0 or t[a] > t[b] or (t[a] == t[b] and a < b)


def seeds_of_round(row_ptr, col_idx, t, elig, owner, seed_dist=2):
    """Unassigned eligible rows that precede every unassigned eligible row
    within distance 2 (neighbours: the columns of a row's entries), or
    within distance 1 when seed_dist is 1 (rule 4)."""
    n = len(row_ptr) - 1
    free = [elig[i] and owner[i] < 0 for i in range(n)]
    m1 = [-1] * n
    for i in range(n):
        best = i if free[i] else -1
        for o in range(row_ptr[i], row_ptr[i + 1]):
            c = col_idx[o]
            if c != i and free[c] and _better(c, best, t):
                best = c
        m1[i] = best
    seeds = []
    for i in range(n):
        if not free[i]:
            continue
        best = m1[i]
        for o in range(row_ptr[i], row_ptr[i + 1]):
            if seed_dist == 1:
                break
            c = col_idx[o]
            if c != i and m1[c] >= 0 and _better(m1[c], best, t):
                best = m1[c]
        if best == i:
            seeds.append(i)
    return seeds


def growth_bids(row_ptr, col_idx, s, t, elig, owner, size):
    """One growth step's choices: {block: row}.  Every unassigned eligible
    row j scores each block with room (< 8 rows) that owns one of its
    neighbours by (sum of s over j's entries into the block, in CSR order) /
    t_j (0 when t_j = 0), and bids for the best (ties: the smaller block
    id); each block takes the bid with the largest FP32-rounded score
    (ties: the smaller row)."""
    n = len(row_ptr) - 1
    best_of_block = {}
    for j in range(n):
        if not elig[j] or owner[j] >= 0:
            continue
        sums = {}
        for o in range(row_ptr[j], row_ptr[j + 1]):
            c = col_idx[o]
            if c == j:
                continue
            B = owner[c]
            if B < 0 or size[B] >= 8:
                continue
            sums[B] = sums[B] + s[o] if B in sums else s[o]
        if not sums:
            continue
        bB, bf = -1, -1.0
        for B, w in sums.items():
            f = w / t[j] if t[j] > 0 else 0.0
            if f > bf or (f == bf and B < bB):
                bB, bf = B, f
        key = (_f32_bits(bf), -j)
        if bB not in best_of_block or key > best_of_block[bB][0]:
            best_of_block[bB] = (key, j)
    return {B: kj[1] for B, kj in best_of_block.items()}


def _rounds(row_ptr, col_idx, s, t, elig, owner, size, members, log=None, seed_dist=2):
    n = len(row_ptr) - 1
    while any(elig[i] and owner[i] < 0 for i in range(n)):
        for i in seeds_of_round(row_ptr, col_idx, t, elig, owner, seed_dist):
            owner[i] = i
            size[i] = 1
            members[i] = [i]
        for _ in range(GROW_STEPS):
            take = growth_bids(row_ptr, col_idx, s, t, elig, owner, size)
            if log is not None:
                log.append((list(owner), dict(take)))    # the owner map the step chose from
            for B, j in take.items():
                owner[j] = B
                size[B] += 1
                members[B].append(j)


def parallel_blocking(row_ptr, col_idx, values, log=None, seed_dist=2):
    """Rule 3: blocks as a list of row lists (ascending rows, ordered by
    first row).
      rounds: seeds = unassigned rows that precede every unassigned row
              within distance 2 in the rule-2 order (larger t_i, then
              smaller row), each a block named by its seed row; then
              GROW_STEPS growth steps (growth_bids), all blocks at once;
              rounds repeat until every row has a block;
      regroup: the rows of blocks with fewer than 8 rows are pooled and
              the rounds rerun on the pool;
      pack:   the blocks still smaller than 8 (pieces), rows ascending,
              in order of their smallest row, cut into consecutive runs of
              PACK_SEGMENT pieces; within a run each piece goes into the most
              recent of the run's last PACK_WINDOW packed blocks with room,
              else a new one (packed blocks never mix runs).
    seed_dist=1 is rule 4: seeds need only precede their distance-1
    neighbours (the multilevel hierarchy's coarse levels)."""
    n = len(row_ptr) - 1
    if n == 0:
        return []
    s = strengths(row_ptr, col_idx, values)
    t = totals(row_ptr, col_idx, s)
    owner, size, members = [-1] * n, [0] * n, [None] * n
    _rounds(row_ptr, col_idx, s, t, [True] * n, owner, size, members, log, seed_dist)
    pool = [size[owner[r]] < 8 for r in range(n)]
    for r in range(n):
        if pool[r]:
            owner[r] = -1
    for B in range(n):
        if size[B] < 8:
            size[B] = 0
    _rounds(row_ptr, col_idx, s, t, pool, owner, size, members, log, seed_dist)
    full = [sorted(members[B]) for B in range(n) if size[B] == 8]
    pieces = sorted((sorted(members[B]) for B in range(n) if 0 < size[B] < 8), key=lambda b: b[0])
    packed = []
    for s0 in range(0, len(pieces), PACK_SEGMENT):
        run = []
        for p in pieces[s0:s0 + PACK_SEGMENT]:
            for q in reversed(run[-PACK_WINDOW:]):
                if len(q) + len(p) <= 8:
                    q.extend(p)
                    break
            else:
                run.append(list(p))
        packed.extend(run)
    out = full + [sorted(q) for q in packed]
    out.sort(key=lambda b: b[0])
    return out
