"""fastsvd and fastcor on the GPU (through the C ABI) against the FP64 oracle,
the golden fixtures of the compiled reference core and the reference's own
tests (tests/testthat/test-rsvd-native-recovery.R,
tests/testthat/test-native-statistical-helpers.R:1-36).

fastsvd (audited_rsvd.hpp:298-414) at precision 64 takes the same attempt and
recovery decisions as the oracle (the audit integers are equal) and agrees
to 1e-9 relative in σ, to 1e-8 in the projector U Uᵀ and in sign-aligned V;
the residual and omitted-direction ratio agree to 1e-5 relative (they are
quotients of small differences). At precision 32 (FP32 operator, FP64 small
math) σ agrees to 1e-5 and the projector to 1e-4 against the oracle on the
same FP32 values, on inputs certified on the first attempt.

fastcor (src/r_api.cpp:4054-4142) agrees with the oracle to 1e-12 absolute at
precision 64 and 2e-6 at precision 32; a constant or non-finite unit gives a
NaN row / column exactly where the oracle has one.
"""
import numpy as np
import pytest

from conftest import align_columns, load_golden, rel
from oracle.oracle import AUDIT_FIELDS, latent_data

pytestmark = pytest.mark.gpu

AUDIT_INTS = ("certified", "deterministic_fallback", "attempts", "effective_oversample",
              "effective_power", "effective_seed")


def _check_same(r, ref, left_only=False, s_tol=1e-9, u_tol=1e-8, audit_tol=1e-5):
    for f in AUDIT_INTS:
        assert r["audit"][f] == ref["audit"][f], (f, r["audit"][f], ref["audit"][f])
    np.testing.assert_allclose(r["s"], ref["s"], rtol=s_tol)
    assert rel(r["U"] @ r["U"].T, ref["U"] @ ref["U"].T) < u_tol
    if not left_only:
        assert r["Vt"] is not None
        assert rel(align_columns(r["Vt"].T, ref["Vt"].T), ref["Vt"].T) < u_tol
    else:
        assert r["Vt"] is None
    for f in ("triplet_residual", "omitted_direction_ratio", "subspace_error",
              "singular_value_error"):
        assert r["audit"][f] == pytest.approx(ref["audit"][f], rel=audit_tol, abs=1e-13), f


@pytest.mark.parametrize("name", ["audited_lowrank", "audited_flat", "audited_tiny",
                                  "audited_wide", "audited_left"])
def test_fastsvd_matches_reference_core(hip, name):
    """Golden fixtures of the compiled reference core: certified at once, after
    three attempts (consensus), the exact route, a wide matrix, left only."""
    fp = hip[0]
    g = load_golden(name)
    left = bool(g["left_only"])
    r = fp.fastsvd(g["A"], int(g["k"]), oversample=int(g["oversample"]), power=int(g["power"]),
                   seed=int(g["seed"]), left_only=left)
    ref = dict(U=g["U"], s=g["s"], Vt=None if left else g["Vt"],
               audit=dict(zip(AUDIT_FIELDS, g["audit"].tolist())))
    _check_same(r, ref, left_only=left)


@pytest.mark.parametrize("shape,k,os_,pw,seed", [((3000, 400), 10, 32, 5, 1),
                                                 ((500, 2500), 6, 10, 2, 4),
                                                 ((4000, 70), 5, 64, 1, 11)])
def test_fastsvd_matches_oracle(hip, oracle, shape, k, os_, pw, seed):
    """Larger operators: several 64-column product chunks, Householder QR for
    sketches wider than 64 columns, a wide matrix."""
    fp = hip[0]
    X = latent_data(shape[0], shape[1], 1, factors=12, noise=0.3, seed=seed)[0]
    r = fp.fastsvd(X, k, oversample=os_, power=pw, seed=seed)
    ref = oracle.audited_rsvd(X, k, oversample=os_, power=pw, seed=seed)
    _check_same(r, ref)


@pytest.mark.parametrize("seed", [1, 7])
@pytest.mark.parametrize("precision", [64, 32])
def test_fastsvd_recovery_slow_spectrum(hip, oracle, seed, precision):
    """tests/testthat/test-rsvd-native-recovery.R: σ = 1 - 0.005 k with no
    oversampling and no power steps needs recovery widenings; the result is
    certified, σ matches to 2e-5 and every triplet residual is at rounding
    level (2e-12 in FP64, 3·2⁻²³·max(m, n) in FP32)."""
    fp = hip[0]
    rng = np.random.default_rng(614)
    U0 = np.linalg.qr(rng.standard_normal((160, 100)))[0]
    V0 = np.linalg.qr(rng.standard_normal((125, 100)))[0]
    sv = 1 - 0.005 * np.arange(100)
    A = (U0 * sv) @ V0.T
    X = A.astype(np.float32) if precision == 32 else A
    r = fp.fastsvd(X, 6, oversample=0, power=0, seed=seed, precision=precision)
    a = r["audit"]
    assert a["certified"] == 1 and a["deterministic_fallback"] == 0
    assert a["attempts"] > 3 and a["effective_oversample"] > 48 and a["effective_power"] >= 6
    assert a["effective_seed"] != seed
    assert a["triplet_residual"] <= 1e-2 and a["omitted_direction_ratio"] <= 1.01
    np.testing.assert_allclose(r["s"], sv[:6], rtol=2e-5)
    T = X.astype(np.float64)
    V = r["Vt"].T
    res = [max(np.linalg.norm(T @ V[:, j] - r["s"][j] * r["U"][:, j]),
               np.linalg.norm(T.T @ r["U"][:, j] - r["s"][j] * V[:, j])) / r["s"][0]
           for j in range(6)]
    tol = 3 * 2.0 ** -23 * max(T.shape) if precision == 32 else 2e-12
    assert max(res) <= tol
    if precision == 64:
        _check_same(r, oracle.audited_rsvd(A, 6, oversample=0, power=0, seed=seed))


def test_fastsvd_fp32_matches_oracle(hip, oracle):
    fp = hip[0]
    X = latent_data(2500, 300, 1, factors=10, noise=0.2, seed=21)[0].astype(np.float32)
    r = fp.fastsvd(X, 8, oversample=16, power=2, seed=3, precision=32)
    ref = oracle.audited_rsvd(X.astype(np.float64), 8, oversample=16, power=2, seed=3)
    for f in AUDIT_INTS:
        assert r["audit"][f] == ref["audit"][f], f
    np.testing.assert_allclose(r["s"], ref["s"], rtol=1e-5)
    assert rel(r["U"] @ r["U"].T, ref["U"] @ ref["U"].T) < 1e-4
    assert rel(align_columns(r["Vt"].T, ref["Vt"].T), ref["Vt"].T) < 1e-4


@pytest.mark.parametrize("shape", [(5, 300), (4000, 4), (5, 5)])
def test_fastsvd_exact_route(hip, shape):
    """min(m, n) < 6: the exact SVD, flagged as the deterministic fallback
    (audited_rsvd.hpp:312-322), equal to the dense SVD."""
    fp = hip[0]
    rng = np.random.default_rng(shape[0] + shape[1])
    X = rng.standard_normal(shape)
    k = 3
    r = fp.fastsvd(X, k, oversample=7, power=2, seed=5)
    U, s, Vt = np.linalg.svd(X, full_matrices=False)
    np.testing.assert_allclose(r["s"], s[:k], rtol=1e-12)
    assert rel(np.abs(r["U"].T @ U[:, :k]), np.eye(k)) < 1e-10
    assert rel(align_columns(r["Vt"].T, Vt[:k].T), Vt[:k].T) < 1e-10
    a = r["audit"]
    assert (a["certified"], a["deterministic_fallback"], a["attempts"]) == (1, 1, 0)
    assert (a["effective_oversample"], a["effective_power"], a["effective_seed"]) == (7, 2, 5)


def test_fastsvd_errors(hip):
    fp = hip[0]
    X = np.random.default_rng(3).standard_normal((50, 20))
    with pytest.raises(fp.HipError, match="controls are invalid"):
        fp.fastsvd(X, 3, oversample=-1)
    with pytest.raises(fp.HipError, match="controls are invalid"):
        fp.fastsvd(X, 0)
    with pytest.raises(ValueError):
        fp.fastsvd(X, 3, precision=16)


def _units(rng, rows, cols):
    return rng.standard_normal((rows, cols)) * rng.uniform(0.5, 3.0, cols) + rng.uniform(-2, 2, cols)


@pytest.mark.parametrize("precision", [64, 32])
def test_fastcor_matches_oracle(hip, oracle, precision):
    fp = hip[0]
    rng = np.random.default_rng(1201)
    tol = 1e-12 if precision == 64 else 2e-6
    dt = np.float64 if precision == 64 else np.float32
    for rows, cols, cols_b in ((12, 7, 7), (700, 90, 130), (3000, 5, 70)):
        a = _units(rng, rows, cols).astype(dt)
        b = _units(rng, rows, cols_b).astype(dt)
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        # columns as units
        np.testing.assert_allclose(fp.fastcor(a, by_row=False, precision=precision),
                                   oracle.fastcor(a64), atol=tol)
        np.testing.assert_allclose(fp.fastcor(a, b, by_row=False, diagonal=False,
                                              precision=precision),
                                   oracle.fastcor(a64, b64), atol=tol)
        # rows as units (the reference default)
        at, bt = np.ascontiguousarray(a.T), np.ascontiguousarray(b.T)
        np.testing.assert_allclose(fp.fastcor(at, bt, by_row=True, diagonal=False,
                                              precision=precision),
                                   oracle.fastcor(at.astype(np.float64), bt.astype(np.float64),
                                                  by_row=True), atol=tol)
        c = _units(rng, cols, rows).astype(dt)
        np.testing.assert_allclose(fp.fastcor(at, c, precision=precision),
                                   oracle.fastcor(at.astype(np.float64), c.astype(np.float64),
                                                  by_row=True, diagonal=True), atol=tol)


def test_fastcor_reference_cases(hip):
    """tests/testthat/test-native-statistical-helpers.R:1-36 (R's defaults
    byrow = TRUE, diag = TRUE) against its closed-form reference."""
    fp = hip[0]
    rng = np.random.default_rng(1201)
    a = rng.standard_normal((12, 7))
    b = rng.standard_normal((12, 7))

    def reference(left, right=None, byrow=True, diagonal=True):
        if not byrow:
            left = left.T
        left = left - left.mean(axis=1, keepdims=True)
        left = left / np.sqrt((left * left).sum(axis=1, keepdims=True))
        if right is None:
            return left @ left.T
        if not byrow:
            right = right.T
        right = right - right.mean(axis=1, keepdims=True)
        right = right / np.sqrt((right * right).sum(axis=1, keepdims=True))
        return (left * right).sum(axis=1) if diagonal else left @ right.T

    np.testing.assert_allclose(fp.fastcor(a), reference(a), atol=1e-12)
    np.testing.assert_allclose(fp.fastcor(a, b), reference(a, b), atol=1e-12)
    np.testing.assert_allclose(fp.fastcor(a, b, diagonal=False), reference(a, b, diagonal=False),
                               atol=1e-12)
    np.testing.assert_allclose(fp.fastcor(a, by_row=False), reference(a, byrow=False), atol=1e-12)


@pytest.mark.parametrize("precision", [64, 32])
def test_fastcor_nan_units(hip, oracle, precision):
    fp = hip[0]
    rng = np.random.default_rng(77)
    dt = np.float64 if precision == 64 else np.float32
    a = _units(rng, 200, 40).astype(dt)
    a[:, 3] = 2.5          # zero norm
    a[17, 9] = np.nan      # non-finite value
    a[5, 30] = np.inf
    got = fp.fastcor(a, by_row=False, precision=precision)
    ref = oracle.fastcor(a.astype(np.float64))
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    np.testing.assert_allclose(got[ok], ref[ok], atol=1e-12 if precision == 64 else 2e-6)
    d = fp.fastcor(np.ascontiguousarray(a.T), np.ascontiguousarray(a.T)[::-1].copy(),
                   precision=precision)
    dref = oracle.fastcor(np.ascontiguousarray(a.T).astype(np.float64),
                          np.ascontiguousarray(a.T)[::-1].astype(np.float64), by_row=True,
                          diagonal=True)
    np.testing.assert_array_equal(np.isnan(d), np.isnan(dref))


def test_fastcor_errors(hip):
    fp = hip[0]
    a = np.ones((5, 4))
    with pytest.raises(fp.HipError, match="same number of correlated features"):
        fp.fastcor(a, np.ones((6, 4)), by_row=False)
    with pytest.raises(fp.HipError, match="matching numbers of rows or columns"):
        fp.fastcor(a, np.ones((5, 3)), by_row=False, diagonal=True)
