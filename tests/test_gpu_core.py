"""The FP32 core hooks (fastpls_hip_core_*_f32), counterparts of the
reference's cuda_core_gemm_f32 / _workspace_create_f32 / _gemm_into_f32 /
_self_gram_into_f32 (src/accelerator_core_backend.h:10-36,
src/core_cuda_backend.cu:38-293), on the GPU through the C ABI.

Checked against numpy FP64 products of the same FP32 values: every
transpose case of op(A) op(B), both self-Gram forms (A Aᵀ and Aᵀ A) with the
lower triangle only and mirrored, on host views whose leading dimension
exceeds their rows, at widths that take the VALU kernels (< 8 output
columns), the matrix cores and several 64-column chunks, all through one
workspace that grows and shrinks between calls. Tolerance: relative
Frobenius error 2e-6 (FP32-accurate products rounded once to FP32; an FP32
result alone carries 6e-8 per element). The reference's contracts are
checked too: an empty inner dimension leaves the output untouched, and
non-conformable shapes fail with its message.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

TOL = 2e-6


def view(rng, rows, cols, pad=5):
    """A rows x cols float32 view of a larger Fortran array (ld = rows + pad)."""
    big = np.asfortranarray(rng.standard_normal((rows + pad, cols + 2)).astype(np.float32))
    return big[:rows, 1:cols + 1]


SHAPES = [(37, 5, 3), (300, 120, 7), (513, 200, 40), (150, 1000, 130), (64, 4096, 9)]


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
def test_core_gemm_into_all_transposes(hip, ta, tb):
    fp = hip[0]
    rng = np.random.default_rng(5 + 2 * ta + tb)
    ws = fp.CoreWorkspace()
    for m, k, n in SHAPES:
        A = view(rng, k, m) if ta else view(rng, m, k)
        B = view(rng, n, k) if tb else view(rng, k, n)
        outbig = np.asfortranarray(np.full((m + 3, n), np.nan, dtype=np.float32))
        out = outbig[:m, :]
        ws.gemm_into(A, B, ta, tb, out)
        a64 = A.astype(np.float64).T if ta else A.astype(np.float64)
        b64 = B.astype(np.float64).T if tb else B.astype(np.float64)
        want = a64 @ b64
        assert rel(out, want) < TOL, (m, k, n, ta, tb, rel(out, want))
        assert np.all(np.isnan(outbig[m:, :]))  # rows past the view are untouched
    ws.close()


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("full", [False, True])
def test_core_self_gram(hip, ta, full):
    fp = hip[0]
    rng = np.random.default_rng(11 + 2 * ta + full)
    ws = fp.CoreWorkspace()
    for rows, cols in [(40, 6), (700, 90), (90, 700), (2000, 150)]:
        A = view(rng, rows, cols)
        dim = cols if ta else rows
        out = np.asfortranarray(np.full((dim, dim), np.nan, dtype=np.float32))
        ws.self_gram_into(A, ta, out, full)
        a = A.astype(np.float64)
        g = a.T @ a if ta else a @ a.T
        lower = np.tril(g)
        assert rel(np.tril(out), lower) < TOL, (rows, cols, ta, full)
        upper = np.triu(out, 1)
        if full:
            np.testing.assert_array_equal(out, out.T)  # mirrored exactly
        else:
            assert np.all(upper == 0.0)
    ws.close()


def test_core_gemm_f32_and_contracts(hip):
    fp = hip[0]
    rng = np.random.default_rng(21)
    A, B = view(rng, 80, 33), view(rng, 33, 12)
    got = fp.core_gemm_f32(A, B)
    assert rel(got, A.astype(np.float64) @ B.astype(np.float64)) < TOL
    ws = fp.CoreWorkspace()
    # an empty inner dimension succeeds and leaves the output untouched
    E1 = np.asfortranarray(np.zeros((4, 0), np.float32))
    E2 = np.asfortranarray(np.zeros((0, 3), np.float32))
    out = np.asfortranarray(np.full((4, 3), 7.0, np.float32))
    ws.gemm_into(E1, E2, False, False, out)
    assert np.all(out == 7.0)
    with pytest.raises(fp.HipError, match="not conformable"):
        ws.gemm_into(view(rng, 5, 4), view(rng, 3, 2), False, False,
                     np.asfortranarray(np.zeros((5, 2), np.float32)))
    with pytest.raises(fp.HipError, match="self-Gram dimensions"):
        ws.self_gram_into(view(rng, 5, 4), False, np.asfortranarray(np.zeros((4, 4), np.float32)),
                          True)
    ws.close()
