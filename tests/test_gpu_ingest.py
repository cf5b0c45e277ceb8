"""GPU runs on ingested MatrixMarket datasets (xp_mtx_load -> device matrix).

Bars (written in each test):
  * the reference's two MatrixMarket files (nasa4704 pattern-symmetric,
    sit100 real-symmetric indefinite) and leading subsets of them: the drop-in
    FP64 SpMV within (k+1) eps sum|a x| per row of the oracle's sequential
    CSR SpMV, and the SELL FP64 SpMV (hi+lo values) within
    (k eps + 2^-46) sum|a x|;
  * an SPD dataset written as a symmetric MatrixMarket file with split
    duplicates and explicit zeros ingests to the generator's CSR bit for
    bit; the accurate iteration on b = A x_ref then recovers x_ref: true
    relative residual <= 1e-10 and relative error <= 1e-10 cond-bound
    (||x - x_ref|| <= 1e-6 ||x_ref|| with cond(A) ~ 1e3-1e4).
"""
from __future__ import annotations

import lzma
import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import mtx as M  # noqa: E402

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
GOLDEN = O.ROOT / "tests" / "golden" / "mtx"


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("mtx")
    out = {}
    for name in ("nasa4704", "sit100"):
        p = d / f"{name}.mtx"
        p.write_bytes(lzma.decompress((GOLDEN / f"{name}.mtx.xz").read_bytes()))
        out[name] = p
    return out


@pytest.mark.parametrize("name", ["nasa4704", "sit100"])
@pytest.mark.parametrize("subset", [None, 3000])
def test_spmv_on_ingested_reference_files(gpu, files, name, subset):
    a, info = X.load_mtx(files[name], subset_n=subset)
    print(f"ingest {name} subset={subset} n={a.n} nnz={a.nnz}", flush=True)
    rp, ci, va, b, _, _ = M.convert(files[name], subset_n=subset)
    assert a.row_ptr.tolist() == rp and np.array_equal(a.b, np.asarray(b))
    x = np.random.default_rng(7).standard_normal(a.n)
    yo = np.asarray(M.csr_spmv(rp, ci, va, x.tolist()))
    absrow = np.asarray(M.csr_spmv(rp, ci, np.abs(va).tolist(), np.abs(x).tolist()))
    k = np.diff(a.row_ptr).max() + 1
    yg = np.zeros(a.n)
    X.spmv_csr_kernel(a.row_ptr, a.col_idx, a.values, x, yg, a.n)
    assert np.all(np.abs(yg - yo) <= k * EPS * absrow)
    # device-resident SELL FP64 SpMV
    m = X.Matrix(a, keep_lo=True)
    inf = m.info()
    perm = m.permutation()
    real = perm >= 0
    xs = np.zeros(inf["n_pad"])
    xs[real] = x[perm[real]]
    xd = X.DeviceArray.from_numpy(xs)
    yd = X.DeviceArray(inf["n_pad"], np.float64)
    m.spmv_f64(xd.ptr, yd.ptr)
    X.synchronize()
    y = np.zeros(a.n)
    y[perm[real]] = yd.numpy()[real]
    assert np.all(np.abs(y - yo) <= (k * EPS + 2.0 ** -46) * absrow)


def spd_as_mtx(a, path, rng):
    """Lower triangle, symmetric banner; a third of the entries split into two
    exact halves on separate lines, explicit zero lines sprinkled in."""
    A = a.to_scipy().tocoo()
    lower = A.row >= A.col
    lines = []
    for r, c, v in zip(A.row[lower] + 1, A.col[lower] + 1, A.data[lower]):
        if rng.random() < 0.33:
            lines += [f"{r} {c} {float(v) / 2!r}", f"{r} {c} {float(v) / 2!r}"]
        else:
            lines.append(f"{r} {c} {float(v)!r}")
        if rng.random() < 0.05:
            lines.append(f"{r} {c} 0.0")
    order = rng.permutation(len(lines))
    body = "\n".join(lines[k] for k in order)
    path.write_text(f"%%MatrixMarket matrix coordinate real symmetric\n% test\n{a.n} {a.n} {len(lines)}\n{body}\n")


@pytest.mark.parametrize("which", ["c1", "fem"])
def test_solve_recovers_x_ref(gpu, tmp_path, which):
    a0 = X.gen_diffusion2d(32, 32, 1.0, 20261204) if which == "c1" else X.gen_fem2d(66, 1.0, 6, 1e3, 0.3, 3)
    p = tmp_path / "a.mtx"
    spd_as_mtx(a0, p, np.random.default_rng(5))
    a, info = X.load_mtx(p)
    print(f"ingested {which}: n={a.n} nnz={a.nnz} entries={info['entries']}", flush=True)
    assert np.array_equal(a.row_ptr, a0.row_ptr) and np.array_equal(a.col_idx, a0.col_idx)
    assert np.array_equal(a.values, a0.values)
    xref = X.reference_solution(a.n)
    np.testing.assert_array_equal(a.b, np.asarray(M.csr_spmv(a.row_ptr.tolist(), a.col_idx.tolist(),
                                                               a.values.tolist(), xref.tolist())))
    tol = 1e-10
    m = X.Matrix(a, keep_lo=True)
    s = X.Pcg(m, rel_tol=tol, accurate=2)
    s.set_rhs(a.b, a.x0)
    rel, repl, iters = s.run_reliable(tol, 0.1)
    x = s.x()
    true = O.residual_norm(a, x) / np.linalg.norm(a.b)
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    print(f"{which}: iters {iters} replacements {repl} true rel {true:.3e} error {err:.3e}", flush=True)
    assert true <= tol * (1 + 1e-6)
    assert err <= 1e-6
