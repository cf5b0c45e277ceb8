"""Closed-form sparse SPD dataset used by the reference's known-answer runs.

Restates the generator algorithm of scripts/generate_cg_dataset.py (reference
file; read, never imported or run): a rows x cols mesh (cols =
ceil(sqrt(size * aspect)), rows = ceil(size / cols)), nodes numbered row-major,
edges to the right and down neighbour plus two sparse families of diagonal
"contact" edges, per-edge weights from fixed trigonometric formulas, each
edge contributing -w off-diagonal and +w to both diagonals, a per-node mass
term on the diagonal, columns sorted per row (generate_cg_dataset.py:11-92);
x_ref and x0 from fixed formulas (:95-114); b = A x_ref in CSR order
(:117-124, :342-346); arrays written with %.17g / str (:127-132).

The reference publishes known answers for size=512, aspect=1.6 (the 18x29
mesh) in its test logs; tests/test_oracle.py checks the oracle against them.

Usage: python gen_cg_dataset.py OUT_DIR [--size 512] [--aspect-ratio 1.6]
"""
from __future__ import annotations

import argparse
import math
from pathlib import Path


def grid_shape(size: int, aspect: float) -> tuple[int, int]:
    cols = max(1, math.ceil(math.sqrt(size * aspect)))
    return max(1, math.ceil(size / cols)), cols


def _w_right(r: int, c: int) -> float:
    base = 1.4 + 0.18 * math.sin(0.17 * r) + 0.11 * math.cos(0.09 * c)
    return max(0.35, base + (0.22 if (r // 9) % 2 == 0 else -0.08))


def _w_down(r: int, c: int) -> float:
    base = 0.95 + 0.15 * math.cos(0.13 * r) + 0.14 * math.sin(0.11 * c)
    return max(0.25, base + (0.20 if (c // 11) % 3 == 1 else 0.02))


def _w_contact(r: int, c: int) -> float:
    return 0.08 + 0.03 * (1.0 + math.sin(0.23 * r + 0.19 * c))


def _mass(i: int, r: int, c: int) -> float:
    return 0.18 + 0.04 * math.sin(0.07 * i) + 0.03 * math.cos(0.15 * r - 0.08 * c)


def build(size: int, aspect: float):
    """Returns (rows, cols, row_ptr, col_idx, values, b, x0) as python lists."""
    rows, cols = grid_shape(size, aspect)
    off: list[dict[int, float]] = [{} for _ in range(size)]
    diag = [0.0] * size

    def link(a: int, b: int, w: float) -> None:
        if a == b or w <= 0.0:
            return
        off[a][b] = off[a].get(b, 0.0) - w
        off[b][a] = off[b].get(a, 0.0) - w
        diag[a] += w
        diag[b] += w

    for i in range(size):
        r, c = divmod(i, cols)
        if c + 1 < cols and i + 1 < size:
            link(i, i + 1, _w_right(r, c))
        if i + cols < size:
            link(i, i + cols, _w_down(r, c))
        if r % 13 == 5 and c % 7 == 2 and c + 1 < cols and i + 2 * cols + 1 < size:
            link(i, i + 2 * cols + 1, _w_contact(r, c))
        if r % 10 == 3 and c % 9 == 4 and c > 0 and i + cols - 1 < size:
            link(i, i + cols - 1, 0.85 * _w_contact(r, c))
    for i in range(size):
        r, c = divmod(i, cols)
        diag[i] += _mass(i, r, c)

    row_ptr = [0]
    col_idx: list[int] = []
    values: list[float] = []
    for i in range(size):
        ent = off[i]
        ent[i] = diag[i]
        for j in sorted(ent):
            col_idx.append(j)
            values.append(ent[j])
        row_ptr.append(len(col_idx))

    centre = max(1, size - 1) * 0.63
    width = max(8.0, size * 0.07)
    x_ref = []
    x0 = []
    for i in range(size):
        r, c = divmod(i, cols)
        x_ref.append(0.75 + 0.18 * math.sin(0.06 * r) + 0.12 * math.cos(0.08 * c)
                     + 0.0009 * i + 0.28 * math.exp(-((i - centre) / width) ** 2))
        x0.append(0.09 * math.cos(0.04 * i) - 0.06 * math.sin(0.09 * r) + 0.03 * math.cos(0.21 * c))
    b = []
    for i in range(size):
        acc = 0.0
        for o in range(row_ptr[i], row_ptr[i + 1]):
            acc += values[o] * x_ref[col_idx[o]]
        b.append(acc)
    return rows, cols, row_ptr, col_idx, values, b, x0


def write_dataset(out_dir: Path, size: int = 512, aspect: float = 1.6) -> None:
    _, _, row_ptr, col_idx, values, b, x0 = build(size, aspect)
    out_dir.mkdir(parents=True, exist_ok=True)
    ints = lambda xs: " ".join(str(v) for v in xs) + "\n"
    flts = lambda xs: " ".join(f"{v:.17g}" for v in xs) + "\n"
    (out_dir / "row_ptr.txt").write_text(ints(row_ptr))
    (out_dir / "col_idx.txt").write_text(ints(col_idx))
    (out_dir / "values.txt").write_text(flts(values))
    (out_dir / "b.txt").write_text(flts(b))
    (out_dir / "x0.txt").write_text(flts(x0))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("out_dir", type=Path)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--aspect-ratio", type=float, default=1.6)
    a = ap.parse_args()
    write_dataset(a.out_dir, a.size, a.aspect_ratio)
