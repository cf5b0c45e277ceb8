"""MatrixMarket ingestion and the binary dataset cache (CPU, no GPU).

The oracle (oracle/mtx.py, the converter's rules restated) is pinned by
scipy.io.mmread -- an independent MatrixMarket reader -- on the reference's
two MatrixMarket files (tests/golden/mtx, xz copies) and on generated files,
by a dense brute-force accumulation of random entry lists with duplicates,
explicit zeros and cancellations, and by subset commutation; the library
(xp_mtx_load and friends) must then equal the oracle bit for bit: CSR
arrays, values and b = A x_ref.
"""
from __future__ import annotations

import lzma
import os
import stat
import sys
import time

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import mtx as M  # noqa: E402

GOLDEN = O.ROOT / "tests" / "golden" / "mtx"


@pytest.fixture(scope="module")
def fixtures(tmp_path_factory):
    d = tmp_path_factory.mktemp("mtx")
    out = {}
    for name in ("nasa4704", "sit100"):
        p = d / f"{name}.mtx"
        p.write_bytes(lzma.decompress((GOLDEN / f"{name}.mtx.xz").read_bytes()))
        out[name] = p
    return out


def write(path, text):
    path.write_text(text)
    return path


def assert_same(a, ref):
    rp, ci, va, b, x0, _ = ref
    assert a.row_ptr.tolist() == rp
    assert a.col_idx.tolist() == ci
    assert np.array_equal(a.values, np.asarray(va, np.float64))
    assert np.array_equal(a.b, np.asarray(b, np.float64))
    assert np.array_equal(a.x0, np.asarray(x0, np.float64))


def scipy_csr(path, subset_n=None):
    """Independent reader: scipy.io.mmread (sums duplicates), zeros dropped,
    leading principal submatrix."""
    import scipy.io
    import scipy.sparse as sp
    A = sp.csr_matrix(scipy.io.mmread(str(path)))
    if subset_n:
        A = A[:subset_n, :subset_n]
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.eliminate_zeros()
    A.sort_indices()
    return A


# ------------------------------------------------------------- oracle pinning
@pytest.mark.parametrize("name", ["nasa4704", "sit100"])
@pytest.mark.parametrize("subset", [None, 1000, 4096])
def test_oracle_matches_scipy_on_reference_files(fixtures, name, subset):
    n, m, rp, ci, va, info = M.parse_matrix(fixtures[name], subset)
    A = scipy_csr(fixtures[name], subset)
    assert (n, m) == A.shape
    assert rp == A.indptr.tolist() and ci == A.indices.tolist()
    np.testing.assert_array_equal(np.asarray(va), A.data)
    d = np.abs(A.diagonal())
    assert info["diag_min_abs"] == d[d > 0].min()
    # symmetric storage expands to a symmetric matrix
    assert (A != A.T).nnz == 0


def random_entries(rng, n, m, count, zeros=0.1, dup=0.3):
    ent = []
    for _ in range(count):
        if ent and rng.random() < dup:
            i, j, _v = ent[rng.integers(len(ent))]
        else:
            i, j = int(rng.integers(1, n + 1)), int(rng.integers(1, m + 1))
        v = 0.0 if rng.random() < zeros else float(rng.choice([-1.5, 0.25, 2.0, -0.75]))
        ent.append((i, j, v))
    return ent


def mtx_text(n, m, ent, field="real", symmetry="general", extra=""):
    lines = [f"%%MatrixMarket matrix coordinate {field} {symmetry}", "% generated", f"{n} {m} {len(ent)}"]
    for i, j, v in ent:
        lines.append(f"{i} {j}" if field == "pattern" else f"{i} {j} {v!r}")
    return "\n".join(lines) + "\n" + extra


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_dense_accumulation(tmp_path, seed):
    """Brute force: a dense accumulator with the converter's rules.  The
    values are dyadic, so sums are exact and order-free; cancellations to 0
    happen and must drop the entry."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 40))
    symmetric = seed % 2 == 1
    ent = random_entries(rng, n, n, 6 * n)
    if symmetric:
        ent = [(max(i, j), min(i, j), v) for i, j, v in ent]
    p = write(tmp_path / "a.mtx", mtx_text(n, n, ent, symmetry="symmetric" if symmetric else "general"))
    for subset in (None, n // 2 + 1):
        k = subset or n
        D = np.zeros((n, n))
        seen = np.zeros((n, n), bool)
        for i, j, v in ent:
            for r, c in ([(i - 1, j - 1), (j - 1, i - 1)] if symmetric and i != j else [(i - 1, j - 1)]):
                if v != 0.0:
                    D[r, c] += v
                    seen[r, c] = True
        D, seen = D[:k, :k], seen[:k, :k]
        nn, mm, rp, ci, va, _ = M.parse_matrix(p, subset)
        assert (nn, mm) == (k, k)
        got = np.zeros((k, k))
        for r in range(k):
            cols = ci[rp[r]:rp[r + 1]]
            assert cols == sorted(cols)
            got[r, cols] = va[rp[r]:rp[r + 1]]
        np.testing.assert_array_equal(got, D)
        assert len(ci) == int(((D != 0) & seen).sum())
        assert_same(X.load_mtx(p, subset_n=subset)[0], M.convert(p, subset_n=subset))


@pytest.mark.parametrize("name", ["nasa4704", "sit100"])
def test_oracle_subset_commutes(fixtures, name):
    """Converting the leading k x k block equals cutting the full conversion."""
    _, _, rp, ci, va, _ = M.parse_matrix(fixtures[name])
    A = scipy_csr(fixtures[name])
    for k in (1, 17, 2000):
        _, _, rs, cs, vs, _ = M.parse_matrix(fixtures[name], k)
        B = A[:k, :k]
        B.eliminate_zeros()
        B.sort_indices()
        assert rs == B.indptr.tolist() and cs == B.indices.tolist()
        np.testing.assert_array_equal(np.asarray(vs), B.data)


def test_reference_solution_properties():
    """Properties the formula fixes: x_i minus the Gaussian bump does not
    depend on n; the bump 0.28 exp(-((i - c)/w)^2) has c = 0.63 max(1, n-1)
    and w = max(8, 0.07 n), located by a quadratic fit of log(bump)."""
    a, b = np.asarray(M.reference_solution(400)), np.asarray(M.reference_solution(3000))
    i = np.arange(400.0)
    smooth = 0.75 + 0.18 * np.sin(0.013 * i) + 0.12 * np.cos(0.017 * i) + 9e-7 * i
    for n, x in ((400, a), (3000, b)):
        bump = x[:400] - smooth if n == 400 else x - (0.75 + 0.18 * np.sin(0.013 * np.arange(n))
                                                     + 0.12 * np.cos(0.017 * np.arange(n)) + 9e-7 * np.arange(n))
        idx = np.arange(len(bump))
        keep = bump > 1e-3
        coef = np.polyfit(idx[keep], np.log(bump[keep] / 0.28), 2)   # -(i^2 - 2ci + c^2)/w^2
        w = np.sqrt(-1.0 / coef[0])
        c = coef[1] * w * w / 2.0
        assert abs(c - 0.63 * (n - 1)) < 1e-6 * n
        assert abs(w - max(8.0, 0.07 * n)) < 1e-6 * n
    assert M.reference_solution(1)[0] == pytest.approx(0.75 + 0.12 + 0.28 * np.exp(-(0.63 / 8.0) ** 2), abs=1e-15)


# ------------------------------------------------------------- library == oracle
@pytest.mark.parametrize("name", ["nasa4704", "sit100"])
@pytest.mark.parametrize("subset", [None, 1, 1000, 10000, 10**7])
def test_library_matches_oracle_on_reference_files(fixtures, name, subset):
    a, info = X.load_mtx(fixtures[name], subset_n=subset)
    ref = M.convert(fixtures[name], subset_n=subset)
    assert_same(a, ref)
    assert info["field"] == ref[5]["field"] and info["symmetry"] == ref[5]["symmetry"]
    assert info["diag_min_abs"] == ref[5]["diag_min_abs"]
    assert (info["nrows"], info["ncols"]) == ref[5]["shape"]
    assert info["entries"] == info["nnz_reported"]


def test_reference_solution_bit_exact():
    for n in (0, 1, 2, 7, 1024, 123457):
        np.testing.assert_array_equal(X.reference_solution(n), np.asarray(M.reference_solution(n), np.float64))


def test_library_edge_cases(tmp_path):
    # comments, blank lines, CRLF, upper-case banner words, extra tokens,
    # '+' signs, integer field, duplicate diagonal, cancellation to zero
    text = ("%%MatrixMarket matrix COORDINATE Integer SYMMETRIC\r\n% c\r\n\r\n%\n  3 3 7  \n"
            "1 1 +4\n2 1 -1 extra\n\n% mid comment\n2 2 4\n3 2 -1\n3 3 4\n3 1 5\n3 1 -5\n")
    p = write(tmp_path / "e.mtx", text)
    a, info = X.load_mtx(p)
    assert_same(a, M.convert(p))
    assert info["field"] == "integer" and info["entries"] == 7
    assert a.to_dense().tolist() == [[4, -1, 0], [-1, 4, -1], [0, -1, 4]]
    # pattern field; a subset that cuts mirrored entries
    p = write(tmp_path / "p.mtx", mtx_text(5, 5, [(1, 1, 0), (4, 2, 0), (5, 5, 0), (2, 2, 0)], "pattern", "symmetric"))
    for sub in (None, 2, 3, 4):
        assert_same(X.load_mtx(p, subset_n=sub)[0], M.convert(p, subset_n=sub))
    # a row with no entries, an empty matrix
    p = write(tmp_path / "z.mtx", mtx_text(3, 3, [(1, 1, 1.0), (3, 3, 2.0)]))
    assert_same(X.load_mtx(p)[0], M.convert(p))
    p = write(tmp_path / "e0.mtx", mtx_text(0, 0, []))
    a, _ = X.load_mtx(p)
    assert a.n == 0 and a.nnz == 0
    # entries outside the declared shape are skipped, as the converter does
    p = write(tmp_path / "o.mtx", mtx_text(2, 2, [(1, 1, 1.0), (3, 1, 9.0), (2, 2, 1.0)]))
    assert_same(X.load_mtx(p)[0], M.convert(p))
    # exponent forms and values that only round-trip with correct rounding
    vals = [1e-300, 5e-324, 1.7976931348623157e308, 0.1, 2.0 / 3.0, -1.25e-17]
    ent = [(k + 1, k + 1, v) for k, v in enumerate(vals)]
    p = write(tmp_path / "f.mtx", mtx_text(6, 6, ent))
    a, _ = X.load_mtx(p)
    assert a.values.tolist() == vals


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "expected coordinate MatrixMarket matrix"),
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "unsupported MatrixMarket field 'complex'"),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n",
     "unsupported MatrixMarket symmetry 'skew-symmetric'"),
    ("%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n", "unsupported MatrixMarket header"),
    ("%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", "unsupported MatrixMarket header"),
    ("%%MatrixMarket matrix coordinate real general\n% only comments\n", "unexpected end of MatrixMarket file"),
    ("%%MatrixMarket matrix coordinate real general\n3 2 1\n1 1 1\n", "is not square after conversion: 3x2"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0D+00\n", "malformed MatrixMarket entry"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n", "malformed MatrixMarket entry"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n", "index below 1"),
])
def test_library_rejects(tmp_path, text, msg):
    p = write(tmp_path / "bad.mtx", text)
    with pytest.raises(X.XplusError, match=msg):
        X.load_mtx(p)
    with pytest.raises((ValueError, IndexError)):
        M.convert(p)


def test_vectors(tmp_path):
    # array storage, several tokens per line
    p = write(tmp_path / "b.mtx", "%%MatrixMarket matrix array real general\n% c\n5 1\n1.5 2\n\n-3\n4e2 5\n")
    np.testing.assert_array_equal(X.read_mtx_vector(p, 5), M.parse_vector(p, 5))
    np.testing.assert_array_equal(X.read_mtx_vector(p, 5, 3), [1.5, 2.0, -3.0])
    # coordinate storage: column 1 only, later lines win, missing rows are 0
    p = write(tmp_path / "c.mtx", "%%MatrixMarket matrix coordinate real general\n4 2 5\n"
              "1 1 7\n2 2 9\n3 1 1\n3 1 2\n9 1 5\n")
    np.testing.assert_array_equal(X.read_mtx_vector(p, 4), [7.0, 0.0, 2.0, 0.0])
    np.testing.assert_array_equal(X.read_mtx_vector(p, 4), M.parse_vector(p, 4))
    p = write(tmp_path / "d.mtx", "%%MatrixMarket matrix array real general\n3 1\n1 2\n")
    with pytest.raises(X.XplusError, match="not enough vector entries"):
        X.read_mtx_vector(p, 3)
    with pytest.raises(X.XplusError, match="vector dimensions do not match dataset"):
        X.read_mtx_vector(p, 4)
    # a dataset with its own b file
    A = write(tmp_path / "A.mtx", mtx_text(3, 3, [(1, 1, 2.0), (2, 2, 2.0), (3, 3, 2.0), (2, 1, 1.0)],
                                           symmetry="symmetric"))
    b = write(tmp_path / "Ab.mtx", "%%MatrixMarket matrix array real general\n3 1\n1\n2\n3\n")
    a, _ = X.load_mtx(A, b_path=b)
    assert a.b.tolist() == [1.0, 2.0, 3.0]
    assert_same(a, M.convert(A, b_path=b))
    a, _ = X.load_mtx(A, b_path=b, subset_n=2)
    assert a.b.tolist() == [1.0, 2.0]


def test_convert_writes_reference_dataset(tmp_path, fixtures):
    meta = X.convert_mtx(fixtures["sit100"], tmp_path / "sit100_n2048", subset_n=2048)
    d = X.load_dataset(tmp_path / "sit100_n2048")
    assert_same(d, M.convert(fixtures["sit100"], subset_n=2048))
    assert meta["shape"] == [2048, 2048] and meta["subset_type"] == "leading_principal_submatrix"
    assert meta["b_source"] == X.X_REF_SOURCE and meta["x0"] == "zeros"
    txt = (tmp_path / "sit100_n2048" / "meta.txt").read_text().splitlines()
    assert txt[:3] == ["name=sit100_n2048", f"source={fixtures['sit100']}", "n=2048"]


# ------------------------------------------------------------- binary cache
def test_binary_round_trip_and_corruption(tmp_path, fixtures):
    a, _ = X.load_mtx(fixtures["nasa4704"])
    p = tmp_path / "a.xpbin"
    a.save_binary(p)
    b = X.load_binary(p)
    for f in ("row_ptr", "col_idx", "values", "b", "x0"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    raw = bytearray(p.read_bytes())
    for off in (200, len(raw) - 3):
        bad = bytearray(raw)
        bad[off] ^= 0x10
        q = tmp_path / "bad.xpbin"
        q.write_bytes(bytes(bad))
        with pytest.raises(X.XplusError, match="checksum mismatch"):
            X.load_binary(q)
    q.write_bytes(bytes(raw[:-8]))
    with pytest.raises(X.XplusError, match="truncated"):
        X.load_binary(q)
    q.write_bytes(b"not a cache" * 20)
    with pytest.raises(X.XplusError, match="not a dataset cache file"):
        X.load_binary(q)
    # an invalid matrix is refused on save, with the text loader's message
    bad = X.Csr(a.row_ptr.copy(), a.col_idx.copy(), a.values, a.b, a.x0)
    bad.col_idx[5] = a.n
    with pytest.raises(X.XplusError, match="col_idx out of range"):
        bad.save_binary(tmp_path / "x.xpbin")


def test_cached_loader(tmp_path):
    d = tmp_path / "ds"
    X.gen_fem2d(30, 1.0, 3, 1e3, 0.3, 5).save(d)
    a0 = X.load_dataset(d)
    a1, hit1 = X.load_dataset_cached(d)
    a2, hit2 = X.load_dataset_cached(d)
    assert (hit1, hit2) == (False, True)
    for f in ("row_ptr", "col_idx", "values", "b", "x0"):
        np.testing.assert_array_equal(getattr(a0, f), getattr(a1, f))
        np.testing.assert_array_equal(getattr(a0, f), getattr(a2, f))
    # a changed text file invalidates the cache
    time.sleep(0.01)
    (d / "b.txt").write_text(" ".join(["1"] * a0.n) + "\n")
    a3, hit3 = X.load_dataset_cached(d)
    assert not hit3 and a3.b.tolist() == [1.0] * a0.n
    assert X.load_dataset_cached(d)[1]
    # a corrupt cache is rebuilt
    c = d / "csr_cache.xpbin"
    raw = bytearray(c.read_bytes())
    raw[-1] ^= 1
    c.write_bytes(bytes(raw))
    a4, hit4 = X.load_dataset_cached(d)
    assert not hit4 and a4.b.tolist() == [1.0] * a0.n
    # the text loader's errors come through unchanged
    (d / "x0.txt").unlink()
    with pytest.raises(X.XplusError, match="failed to open"):
        X.load_dataset_cached(d)


@pytest.mark.skipif(os.geteuid() == 0, reason="root ignores directory permissions")
def test_cached_loader_read_only(tmp_path):
    d = tmp_path / "ro"
    X.gen_diffusion2d(10, 10, 1.0, 1).save(d)
    d.chmod(stat.S_IRUSR | stat.S_IXUSR)
    try:
        a, hit = X.load_dataset_cached(d)
        assert not hit and a.n == 100
    finally:
        d.chmod(stat.S_IRWXU)


def test_cached_loader_outside_dir(tmp_path):
    d = tmp_path / "ds"
    X.gen_diffusion2d(10, 10, 1.0, 1).save(d)
    c = tmp_path / "elsewhere.xpbin"
    assert not X.load_dataset_cached(d, c)[1]
    assert c.exists() and not (d / "csr_cache.xpbin").exists()
    assert X.load_dataset_cached(d, c)[1]


def test_cache_is_faster_than_text(tmp_path):
    d = tmp_path / "big"
    X.gen_diffusion2d(400, 400, 1.0, 3).save(d)       # 160k rows, 0.8M nnz
    t0 = time.perf_counter()
    X.load_dataset(d)
    t_text = time.perf_counter() - t0
    X.load_dataset_cached(d)
    t0 = time.perf_counter()
    _, hit = X.load_dataset_cached(d)
    t_cache = time.perf_counter() - t0
    print(f"text {t_text * 1e3:.1f} ms, cache {t_cache * 1e3:.1f} ms")
    assert hit and t_cache < t_text / 3
