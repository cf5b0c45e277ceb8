"""Batched uneven systems (the 256-system workload's family: 3D 7-point
diffusion, log-normal coefficients sigma 1.5, edges k in 16..64) including
the largest systems (k = 64, n = 262,144), each against its own FP64 golden.

Bars:
  * accurate iteration with residual replacement: every system's FP64 true
    residual (recomputed on the CPU) <= 1e-6 ||b||, x within relative 1e-5
    of the golden's, iterations in [golden - 2, 1.25 golden] (FP32 r and z
    cost iterations: +10-20% measured on the FEM family);
  * FP32 iteration: every system converged on its recursive residual with
    an iteration count within max(2, 3%) of its golden's;
  * marking the workgroups of stopped systems idle between graph launches
    (the default) changes nothing: the same iterations and bit-identical x
    as with every workgroup launched (full_work_map).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

pytestmark = pytest.mark.gpu

KS = (16, 29, 41, 64, 53, 64)


@pytest.fixture(scope="module")
def batch():
    systems = [X.gen_diffusion3d(k, 1.5, 20261203 + 1000 * 4 + i) for i, k in enumerate(KS)]
    goldens = [O.golden(a, tau=(1e-6 * np.linalg.norm(a.b)) ** 2, trace_cap=1) for a in systems]
    return systems, goldens


def split(x, systems):
    out, off = [], 0
    for a in systems:
        out.append(x[off:off + a.n])
        off += a.n
    return out


def test_accurate_batch_with_largest_systems_matches_golden(gpu, batch):
    systems, goldens = batch
    m = X.Matrix(systems, keep_lo=True)
    s = X.Pcg(m, rel_tol=1e-6, accurate=True)
    s.set_rhs(np.concatenate([a.b for a in systems]))
    rel, repl, _ = s.run_reliable(1e-6, 0.0)
    for i, (a, g, xi) in enumerate(zip(systems, goldens, split(s.x(), systems))):
        it = s.status(i)["iterations"]
        res = np.linalg.norm(a.b - a.to_scipy() @ xi) / np.linalg.norm(a.b)
        dx = np.linalg.norm(xi - g["x"]) / np.linalg.norm(g["x"])
        print(f"sys {i} k={KS[i]} n={a.n}: {it} iterations (golden {g['iterations']}), rel {res:.3e}, dx {dx:.2e}",
              flush=True)
        assert res <= 1e-6 and rel[i] <= 1e-6
        assert dx <= 1e-5
        assert g["iterations"] - 2 <= it <= 1.25 * g["iterations"]


def test_fp32_batch_with_largest_systems_matches_golden_counts(gpu, batch):
    systems, goldens = batch
    m = X.Matrix(systems)
    s = X.Pcg(m, rel_tol=1e-6, fused=True)
    s.set_rhs(np.concatenate([a.b for a in systems]))
    s.run()
    for i, g in enumerate(goldens):
        st = s.status(i)
        print(f"sys {i} k={KS[i]}: {st['iterations']} iterations (golden {g['iterations']})", flush=True)
        assert st["status"] == X.CONVERGED
        assert abs(st["iterations"] - g["iterations"]) <= max(2, 0.03 * g["iterations"])


@pytest.mark.parametrize("accurate", [False, True])
def test_idle_workgroups_of_stopped_systems_change_nothing(gpu, batch, accurate):
    systems, _ = batch
    m = X.Matrix(systems, keep_lo=True)
    b = np.concatenate([a.b for a in systems])
    res = []
    for full in (True, False):
        s = X.Pcg(m, rel_tol=1e-6, accurate=accurate, fused=True, full_work_map=full, graph_iters=16)
        s.set_rhs(b)
        if accurate:
            s.run_reliable(1e-6, 0.0)
        else:
            s.run()
        res.append(([s.status(i)["iterations"] for i in range(len(systems))], s.x(), s.last_timing()[0]))
        s.close()
    print(f"accurate={accurate}: iterations {res[0][0]}; {res[0][2]:.2f} ms with every workgroup, "
          f"{res[1][2]:.2f} ms with stopped systems idle", flush=True)
    assert res[0][0] == res[1][0]
    np.testing.assert_array_equal(res[0][1], res[1][1])
