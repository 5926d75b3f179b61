"""The FP64 streaming products at ragged shapes, against numpy FP64.

`fastpls_hip_gemm` at precision 64 runs out = A W through the score-pass
kernels (8..64 right-hand columns on the FP64 matrix cores, fewer on the FP64
VALU kernels, the reduction split over column ranges); FP64 `fastcor` over
column units runs the cross-product kernels (A_sᵀ B_s over the sample rows).
The shapes cover partial 64-row and 256-row tiles, partial 32-column steps,
every matrix-core width (16, 32, 48, 64 columns) and several splits, so the
pipelined kernels' prologue, steady state and ragged last step all run.
Tolerance: 1e-12 relative (FP64 products, FP64 sums in a fixed order).
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,inner", [(1, 1), (63, 31), (64, 32), (65, 33), (257, 100),
                                        (1000, 4097), (4100, 130)])
@pytest.mark.parametrize("cols", [3, 8, 16, 17, 33, 48, 64, 65, 130])
def test_fp64_score_products_at_ragged_shapes(hip, rows, inner, cols):
    fp = hip[0]
    rng = np.random.default_rng(rows * 7919 + inner * 31 + cols)
    A = np.asfortranarray(rng.standard_normal((rows, inner)) + 3.0)
    W = np.asfortranarray(rng.standard_normal((inner, cols)))
    got = fp.gemm(A, W, precision=64)
    assert rel(got, A @ W) < 1e-12


@pytest.mark.parametrize("samples", [65, 130, 1000, 4100])
@pytest.mark.parametrize("ua,ub", [(1, 8), (127, 17), (129, 33), (300, 64), (70, 100)])
def test_fp64_cross_products_at_ragged_shapes(hip, samples, ua, ub):
    fp = hip[0]
    rng = np.random.default_rng(samples * 131 + ua * 17 + ub)
    a = np.asfortranarray(rng.standard_normal((samples, ua)) * 2.0 + 1.0)
    b = np.asfortranarray(rng.standard_normal((samples, ub)) + a[:, :1] * 0.5)
    got = fp.fastcor(a, b, by_row=False, diagonal=False, precision=64)
    ref = np.corrcoef(a.T, b.T)[:ua, ua:]
    assert np.abs(got - ref).max() < 1e-12
