"""CPU oracle (pure Python) of the reference's MatrixMarket -> CSR dataset
conversion.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the product).  It
restates, function by function, scripts/download_suitesparse_data.py:

  parse_matrix      parse_coordinate_matrix (:224-284): coordinate storage;
                    real|integer|pattern field; general|symmetric symmetry;
                    zeros skipped on entry, duplicates summed in file order,
                    the mirror (j, i) of an off-diagonal symmetric entry added
                    when inside the subset, columns ascending, sums equal to 0
                    dropped; info: diag_min_abs / has_nonzero_diag_entries.
  parse_vector      parse_vector (:287-336): array (tokens in order, several
                    per line allowed) or coordinate (column 1 only, last write
                    wins) right-hand sides.
  reference_solution  (:339-349) the smooth x_ref.
  csr_spmv          (:352-358) sequential FP64 row sums.
  convert           convert_dataset (:418-444): b from the b file, else
                    b = A x_ref; x0 = 0; square check.

Deliberate difference, shared with the library: a row or column index below
1 raises ValueError (the converter would wrap index 0 to the last row).

Pinning (tests/test_mtx.py): scipy.io.mmread (an independent MatrixMarket
reader) on the reference's two MatrixMarket files and on generated files;
a dense brute-force accumulation of random entry lists with duplicates,
zeros and cancellations; subset commutation (converting a subset equals the
subset of the full conversion); reference_solution through properties its
formula fixes (size-independent smooth part, Gaussian bump centred at
0.63 (n - 1) with width max(8, 0.07 n)).
"""
from __future__ import annotations

import math


def _lines(path):
    with open(path, "r", encoding="utf-8", errors="ignore") as handle:
        yield from handle


def _header(path):
    it = _lines(path)
    first = next(it, "")
    header = first.strip().split()
    if len(header) < 5 or header[0] != "%%MatrixMarket" or header[1] != "matrix":
        raise ValueError(f"unsupported MatrixMarket header in {path}")
    for raw in it:
        line = raw.strip()
        if line and not line.startswith("%"):
            return header, line, it
    raise ValueError("unexpected end of MatrixMarket file")


def _index(token):
    v = int(token) - 1
    if v < 0:
        raise ValueError("MatrixMarket index below 1")
    return v


def parse_matrix(path, subset_n=None):
    """(n, m, row_ptr, col_idx, values, info) as lists."""
    header, dims, it = _header(path)
    storage, field, symmetry = (h.lower() for h in header[2:5])
    if storage != "coordinate":
        raise ValueError(f"expected coordinate MatrixMarket matrix: {path}")
    if field not in ("real", "integer", "pattern"):
        raise ValueError(f"unsupported MatrixMarket field '{field}' in {path}")
    if symmetry not in ("general", "symmetric"):
        raise ValueError(f"unsupported MatrixMarket symmetry '{symmetry}' in {path}")
    tok = dims.split()
    nrows, ncols = int(tok[0]), int(tok[1])
    int(tok[2])
    rows_out = min(nrows, subset_n) if subset_n else nrows
    cols_out = min(ncols, subset_n) if subset_n else ncols
    acc = [dict() for _ in range(rows_out)]

    def put(r, c, v):
        if v != 0.0:
            acc[r][c] = acc[r].get(c, 0.0) + v

    for raw in it:
        line = raw.strip()
        if not line or line.startswith("%"):
            continue
        parts = line.split()
        i, j = _index(parts[0]), _index(parts[1])
        v = 1.0 if field == "pattern" else float(parts[2])
        if i < rows_out and j < cols_out:
            put(i, j, v)
        if symmetry == "symmetric" and i != j and j < rows_out and i < cols_out:
            put(j, i, v)

    row_ptr, col_idx, values, diag = [0], [], [], []
    for r in range(rows_out):
        for c in sorted(acc[r]):
            v = acc[r][c]
            if v == 0.0:
                continue
            col_idx.append(c)
            values.append(v)
            if c == r:
                diag.append(abs(v))
        row_ptr.append(len(col_idx))
    info = {"field": field, "symmetry": symmetry, "shape": (nrows, ncols),
            "diag_min_abs": min(diag) if diag else None, "has_diag": bool(diag)}
    return rows_out, cols_out, row_ptr, col_idx, values, info


def parse_vector(path, expected_n, subset_n=None):
    header, dims, it = _header(path)
    storage, field = header[2].lower(), header[3].lower()
    if field not in ("real", "integer", "pattern"):
        raise ValueError(f"unsupported vector field '{field}' in {path}")
    limit = min(expected_n, subset_n) if subset_n else expected_n
    out = [0.0] * limit
    if storage == "array":
        nrows, ncols = (int(t) for t in dims.split()[:2])
        if nrows < limit or ncols < 1:
            raise ValueError(f"vector dimensions do not match dataset: {path}")
        k = 0
        for raw in it:
            if k >= limit:
                break
            line = raw.strip()
            if not line or line.startswith("%"):
                continue
            for t in line.split():
                if k >= limit:
                    break
                out[k] = float(t)
                k += 1
        if k < limit:
            raise ValueError(f"not enough vector entries in {path}")
        return out
    if storage == "coordinate":
        nrows, ncols, _ = (int(t) for t in dims.split()[:3])
        if nrows < limit or ncols < 1:
            raise ValueError(f"vector dimensions do not match dataset: {path}")
        for raw in it:
            line = raw.strip()
            if not line or line.startswith("%"):
                continue
            parts = line.split()
            i, j = _index(parts[0]), _index(parts[1])
            if i >= limit or j != 0:
                continue
            out[i] = 1.0 if field == "pattern" else float(parts[2])
        return out
    raise ValueError(f"unsupported vector MatrixMarket storage '{storage}' in {path}")


def reference_solution(n):
    c = max(1, n - 1) * 0.63
    w = max(8.0, n * 0.07)
    return [0.75 + 0.18 * math.sin(0.013 * i) + 0.12 * math.cos(0.017 * i) + 0.0000009 * i
            + 0.28 * math.exp(-((i - c) / w) ** 2) for i in range(n)]


def csr_spmv(row_ptr, col_idx, values, x):
    y = []
    for r in range(len(row_ptr) - 1):
        s = 0.0
        for o in range(row_ptr[r], row_ptr[r + 1]):
            s += values[o] * x[col_idx[o]]
        y.append(s)
    return y


def convert(path, b_path=None, subset_n=None):
    """(row_ptr, col_idx, values, b, x0, info) of one dataset."""
    n, m, rp, ci, va, info = parse_matrix(path, subset_n)
    if n != m:
        raise ValueError(f"not square after conversion: {n}x{m}")
    b = parse_vector(b_path, n, subset_n) if b_path else csr_spmv(rp, ci, va, reference_solution(n))
    return rp, ci, va, b, [0.0] * n, info
