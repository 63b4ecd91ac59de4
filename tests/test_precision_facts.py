"""Arithmetic facts the FP32 path relies on (CPU).

Total depth H = eta*Gamma3 + Depth (etauv_solver.F:373) with Gamma3 = 1 is
the sum of two FP32 numbers: forming it in FP64 and rounding back to FP32
gives the same bits as the FP32 sum (double rounding through 53 bits is
innocuous for a sum of two 24-bit numbers), and near-dry cells (Depth close
to -eta) sum exactly (Sterbenz).  So an FP64 total depth in FP32 runs would
change nothing."""
import numpy as np


def test_fp64_sum_of_two_fp32_rounds_like_fp32():
    rng = np.random.default_rng(11)
    n = 2_000_000
    depth = (rng.random(n) * 20.0 - 2.0).astype(np.float32)
    # near-dry: eta close to -depth, plus generic and small surfaces
    eta = np.concatenate([
        (-depth[: n // 2] * (1 + (rng.random(n // 2) - 0.5) * 1e-3)).astype(np.float32),
        (rng.standard_normal(n - n // 2) * 0.5).astype(np.float32),
    ])
    s32 = eta + depth
    s64 = (eta.astype(np.float64) + depth.astype(np.float64)).astype(np.float32)
    assert s32.dtype == np.float32
    assert np.array_equal(s32.view(np.uint32), s64.view(np.uint32))
    # Sterbenz: where depth/2 <= -eta <= 2*depth the FP32 difference is exact
    m = (depth > 0) & (-eta >= depth / 2) & (-eta <= 2 * depth)
    assert m.sum() > n // 4
    exact = eta[m].astype(np.float64) + depth[m].astype(np.float64)
    assert np.array_equal(s32[m].astype(np.float64), exact)
