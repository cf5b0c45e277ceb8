"""The measurement helpers the bench line is built from (no reference
counterpart): the streaming reference kernels (xp_stream_time, every mode)
and the per-kernel event timing of the accurate iteration, whole and split
by iteration parity (xp_pcg_kernel_times, xp_pcg_kernel_times_split).

Bars (written in each test): every timing is finite and positive; a
streaming mode outside 0..4 is refused; the parity split agrees with the
plain average within the run-to-run spread allowed here (the mean of the
even and odd averages within 25 % of the whole average, measured back to
back); the solver still solves to its tolerance after the timing runs,
which leave its vectors as after fixed iterations.
"""
from __future__ import annotations

import numpy as np
import pytest

import xplus_amd as X

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_stream_modes_time(gpu, mode):
    us = X.stream_time(1 << 22, mode, 256, 5)
    assert np.isfinite(us) and us > 0.0


def test_stream_mode_out_of_range_refused(gpu):
    with pytest.raises(X.XplusError):
        X.stream_time(1 << 20, 5, 256, 1)


def test_kernel_times_split_accurate_blocks(gpu):
    a = X.gen_fem2d(120, 1.0, 20, 1e3, 0.3, 20261977)
    m = X.Matrix(a, keep_lo=True, block_rows=8, blocking=3)
    s = X.Pcg(m, rel_tol=1e-6, accurate=True)
    s.set_rhs(a.b, a.x0)
    s.set_preconditioner(1)
    whole = s.kernel_times(40)
    even, odd = s.kernel_times_split(40)
    assert whole.shape == (2,) and even.shape == (2,) and odd.shape == (2,)
    assert np.all(np.isfinite(whole)) and np.all(whole > 0.0)
    assert np.all(even > 0.0) and np.all(odd > 0.0)
    mean = (even + odd) / 2.0
    assert np.all(np.abs(mean - whole) <= 0.25 * whole), (whole, even, odd)
    rel, _, iters = s.run_reliable(1e-6, 0.0)
    assert rel[0] <= 1e-6 and iters > 0
    s.close()
    m.close()


def test_device_pool_reuses_and_trims(gpu):
    """The caching allocator (no reference counterpart): a solve on a
    matrix whose buffers reuse cached blocks gives the same bits as the first
    one; closing keeps blocks cached; trimming returns them all."""
    a = X.gen_fem2d(200, 1.0, 20, 1e3, 0.3, 20261978)
    outs = []
    for k in range(3):
        m = X.Matrix(a, keep_lo=True, block_rows=8, blocking=3)
        s = X.Pcg(m, rel_tol=1e-6, accurate=True)
        s.set_rhs(a.b, a.x0)
        s.set_preconditioner(1)
        rel, _, iters = s.run_reliable(1e-6, 0.0)
        outs.append((iters, s.x()))
        s.close()
        m.close()
        X.synchronize()
        assert X.pool_bytes() > 0
        if k == 1:
            X.pool_trim()
            assert X.pool_bytes() == 0
    for iters, x in outs[1:]:
        assert iters == outs[0][0]
        np.testing.assert_array_equal(x, outs[0][1])
