"""Edge cases of the solver against the FP64 golden (oracle):

  * zero right-hand side with x0 = 0: the golden stops at iteration 0
    (r^T r = 0 <= tau); so does every iteration form here, with x = 0
    exactly and a true relative residual reported as 0;
  * a diagonal matrix: Jacobi-PCG takes one iteration (z = D^-1 r is the
    exact correction), as the golden does;
  * systems smaller than a 64-row slice and than an 8-row block (n = 1, 2,
    7, 9, 63, 65): FP32 and accurate iterations within +-1 of the golden's
    count and x within relative 1e-5 of its x; with one exact 8x8 block
    covering the whole system (n <= 8) the preconditioner is the inverse up
    to its FP16 factor (relative ~2^-11 per entry): one step for n = 1,
    where the factor is exact, and for n > 1 the count of the FP64 block
    PCG oracle driven by the operator the device stores (+-1), at most 2.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import block_mlp as B  # noqa: E402

pytestmark = pytest.mark.gpu

FORMS = [dict(fused=False), dict(fused=True), dict(accurate=True)]


def rel_err(x, y):
    return np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)


@pytest.mark.parametrize("form", FORMS, ids=["three-kernel", "fused", "accurate"])
def test_zero_rhs_stops_at_iteration_zero(gpu, form):
    a = X.gen_diffusion2d(20, 17, 1.0, 3)
    b = np.zeros(a.n)
    g = O.golden(a, b=b, x0=np.zeros(a.n), tau=1e-10)
    assert g["iterations"] == 0 and g["converged"]
    m = X.Matrix(a, keep_lo=True)
    s = X.Pcg(m, rel_tol=1e-6, **form)
    s.set_rhs(b)
    s.run()
    st = s.status()
    assert st["status"] == X.CONVERGED and st["iterations"] == 0
    assert not s.x().any()
    if form.get("accurate") or form.get("fused"):
        rel, _, iters = s.run_reliable(1e-6, 0.0)
        assert iters == 0 and rel[0] == 0.0 and not s.x().any()
    s.close()
    m.close()


@pytest.mark.parametrize("form", FORMS, ids=["three-kernel", "fused", "accurate"])
def test_diagonal_matrix_takes_one_iteration(gpu, form):
    n = 300
    d = np.random.default_rng(5).uniform(0.5, 4.0, n)
    a = X.Csr(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d,
              np.random.default_rng(6).uniform(-1, 1, n), np.zeros(n))
    g = O.golden(a, tau=(1e-6 * np.linalg.norm(a.b)) ** 2)
    assert g["iterations"] == 1
    m = X.Matrix(a, keep_lo=True)
    s = X.Pcg(m, rel_tol=1e-6, **form)
    s.set_rhs(a.b)
    s.run()
    st = s.status()
    assert st["status"] == X.CONVERGED and st["iterations"] == 1
    assert rel_err(s.x(), g["x"]) <= 1e-6
    s.close()
    m.close()


SMALL = {1: (1, 1), 2: (2, 1), 7: (7, 1), 9: (3, 3), 63: (9, 7), 65: (13, 5)}


@pytest.mark.parametrize("form", FORMS, ids=["three-kernel", "fused", "accurate"])
@pytest.mark.parametrize("n", list(SMALL))
def test_systems_smaller_than_a_slice(gpu, n, form):
    a = X.gen_diffusion2d(*SMALL[n], 1.0, 100 + n)
    assert a.n == n
    tau = (1e-6 * np.linalg.norm(a.b)) ** 2
    g = O.golden(a, tau=tau)
    m = X.Matrix(a, keep_lo=True)
    s = X.Pcg(m, rel_tol=1e-6, **form)
    s.set_rhs(a.b)
    s.run()
    st = s.status()
    print(f"n={n} {form}: {st['iterations']} iterations (golden {g['iterations']})", flush=True)
    assert st["status"] == X.CONVERGED
    assert abs(st["iterations"] - g["iterations"]) <= 1
    assert rel_err(s.x(), g["x"]) <= 1e-5
    s.close()
    m.close()


@pytest.mark.parametrize("blocking", [0, 3])
@pytest.mark.parametrize("n", [1, 2, 7])
def test_one_exact_block_is_the_inverse(gpu, n, blocking):
    a = X.gen_diffusion2d(*SMALL[n], 1.0, 200 + n)
    m = X.Matrix(a, keep_lo=True, block_rows=8, blocking=blocking)
    assert m.num_blocks() == 1
    s = X.Pcg(m, rel_tol=1e-6, accurate=True)
    s.set_preconditioner(1)
    s.set_rhs(a.b)
    rel, _, iters = s.run_reliable(1e-6, 0.0)
    x = s.x()
    op = s.block_operator()
    groups = [[int(r) for r in row if r >= 0] for row in m.block_rows()]
    A = a.to_scipy()
    _, it_o, _ = B.block_pcg(A, a.b, op, (1e-6 * np.linalg.norm(a.b)) ** 2, groups=groups)
    xe = np.linalg.solve(A.toarray(), a.b)
    print(f"n={n} blocking {blocking}: {iters} iterations (FP64 oracle with the stored operator {it_o}), "
          f"rel {rel[0]:.2e}", flush=True)
    assert rel[0] <= 1e-6 and rel_err(x, xe) <= 1e-5
    if n == 1:
        assert iters == 1
    else:
        assert abs(iters - it_o) <= 1 and iters <= 2
    s.close()
    m.close()
