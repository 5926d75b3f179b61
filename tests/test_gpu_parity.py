"""GPU parity through the C ABI against the FP64 oracle and the golden vectors.

Tolerances (relative Frobenius error unless stated):
  * precision 64: 1e-9 on coefficients B and fitted/predicted Ŷ;
  * precision 32 (X stored FP32, FP64 accumulation and small math): 1e-5 on
    B and Ŷ against the oracle run on the same FP32 values widened to FP64;
  * labels: identical except rows whose top-2 margin is below 1e-6.
"""
import numpy as np
import pytest

from conftest import align_columns, load_golden, rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-5}


def coef(m, k):
    return m["R"][:, :k] @ m["Q"][:, :k].T


def hip_coef(model, k):
    R = model.export("R").astype(np.float64)
    Q = model.export("Q").astype(np.float64)
    return R[:, :k] @ Q[:, :k].T


def fit(hip, X, Y=None, labels=None, classes=None, method=0, precision=64, **kw):
    fp = hip[0]
    return fp.ResidentModel.fit(X, Y, labels=labels, classes=classes, method=method,
                                precision=precision, **kw)


# ------------------------------------------------------------------ SIMPLS
@pytest.mark.parametrize("precision", [64, 32])
def test_simpls_golden_small(hip, oracle, precision):
    g = load_golden("simpls_dense_small")
    X = g["X"].astype(np.float32 if precision == 32 else np.float64)
    m = fit(hip, X, g["Y"], precision=precision, components=int(g["ncomp"]),
            oversample=int(g["oversample"]), power=int(g["power"]), seed=int(g["seed"]))
    o = oracle.simpls_fit(X.astype(np.float64), g["Y"].astype(X.dtype).astype(np.float64),
                          ncomp=int(g["ncomp"]), oversample=int(g["oversample"]),
                          power=int(g["power"]), seed=int(g["seed"]))
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(hip_coef(m, k), coef(o, k)) < TOL[precision], k
        if precision == 64:
            assert rel(hip_coef(m, k), coef(g, k)) < 1e-8, k
    np.testing.assert_allclose(m.export("mX"), o["mx"], rtol=1e-6 if precision == 32 else 1e-12,
                               atol=1e-6 if precision == 32 else 1e-12)
    np.testing.assert_allclose(m.export("vX"), o["sx"], rtol=1e-6 if precision == 32 else 1e-12)
    T = m.export("T").astype(np.float64)
    np.testing.assert_allclose(T.T @ T, np.eye(T.shape[1]), atol=1e-6 if precision == 32 else 1e-10)
    m.close()


@pytest.mark.parametrize("name", ["simpls_dense_sketch", "simpls_labels_small"])
def test_simpls_golden_sketch_dependent(hip, oracle, name):
    g = load_golden(name)
    kw = dict(components=int(g["ncomp"]), oversample=int(g["oversample"]), power=int(g["power"]),
              seed=int(g["seed"]))
    if "labels" in g:
        m = fit(hip, g["X"], labels=g["labels"], classes=int(g["classes"]), **kw)
    else:
        m = fit(hip, g["X"], g["Y"], **kw)
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(hip_coef(m, k), coef(g, k)) < 1e-7, k
    R = m.export("R").astype(np.float64)
    assert rel(align_columns(R, g["R"]), g["R"]) < 1e-7
    m.close()


def test_simpls_c1_workload(hip, oracle):
    """n=2000, p=1000, q=3, 10 components, FP64 and FP32 paths."""
    g = load_golden("simpls_c1")
    X, Y = latent_data(2000, 1000, 3, factors=10, noise=0.1, seed=0)
    for precision in (64, 32):
        Xp = X.astype(np.float64 if precision == 64 else np.float32)
        m = fit(hip, Xp, Y, precision=precision, components=10, seed=1)
        ref = g if precision == 64 else oracle.simpls_fit(Xp.astype(np.float64),
                                                          Y.astype(np.float32).astype(np.float64),
                                                          ncomp=10, seed=1)
        for k in (1, 2, 5, 10):
            assert rel(hip_coef(m, k), coef(ref, k)) < TOL[precision], (precision, k)
        # fitted values on the training data, every prefix
        pred = m.predict_path(Xp, list(range(1, 11))).astype(np.float64)
        Xs = (Xp.astype(np.float64) - ref["mx"]) / ref["sx"]
        for k in (1, 5, 10):
            yh = Xs @ coef(ref, k) + ref["my"]
            assert rel(pred[:, :, k - 1], yh) < TOL[precision] * 10, (precision, k)
        m.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_simpls_midsize_refresh(hip, oracle, precision):
    """p=5000, q=14, 8 components (block 1): the refresh's tall CholQR2 of the
    5000 x 11 sketch (matrix-core Grams), the right-hand CholQR2 and the
    projected-Gram eigen in LDS, the row-parallel update, against the oracle
    on the same values."""
    X, Y = latent_data(1500, 5000, 14, factors=12, noise=0.1, seed=21, decay=0.8)
    dt = np.float64 if precision == 64 else np.float32
    Xp, Yp = X.astype(dt), Y.astype(dt)
    m = fit(hip, Xp, Yp, precision=precision, components=8, oversample=10, power=2, seed=3)
    ref = oracle.simpls_fit(Xp.astype(np.float64), Yp.astype(np.float64), ncomp=8, oversample=10,
                            power=2, seed=3)
    for k in (1, 4, 8):
        assert rel(hip_coef(m, k), coef(ref, k)) < TOL[precision], k
    m.close()


def test_simpls_rank_deficient_response(hip, oracle):
    """Y of rank 2 in 12 columns: the refresh sketches (w = 11) have rank 2, so
    both CholQR2 factorisations break down and the Householder QRs take over
    (right-hand in LDS, tall in global memory); the two leading components
    against the oracle."""
    X, Y0 = latent_data(800, 3000, 2, factors=6, noise=0.1, seed=22)
    Y = Y0 @ np.random.default_rng(22).standard_normal((2, 12))
    m = fit(hip, X, Y, components=2, oversample=10, power=2, seed=5)
    ref = oracle.simpls_fit(X, Y, ncomp=2, oversample=10, power=2, seed=5)
    for k in (1, 2):
        assert rel(hip_coef(m, k), coef(ref, k)) < 1e-8, k
    m.close()


def test_simpls_block_routing_classification(hip, oracle):
    """n*p*K >= 5e8 routes classification to a candidate block (simpls.hpp:43-62)."""
    X, lab = class_data(10000, 2500, 20, latent=15, seed=2)
    ctl = oracle.controls(10000, 2500, 20, 15, classification=True)
    assert ctl["block"] == 15 and ctl["batch_geometry"] == 1
    m = fit(hip, X, labels=lab, classes=20, components=15, seed=1)
    assert m.controls()["block"] == 15
    o = oracle.simpls_fit(X, labels=lab, classes=20, ncomp=15, seed=1)
    for k in (1, 8, 15):
        assert rel(hip_coef(m, k), coef(o, k)) < 1e-9, k
    # LDA labels on held-out rows identical except near-ties
    Xt, labt = class_data(3000, 2500, 20, latent=15, seed=2)
    got = m.classify_path(Xt, [15], lda=True, top=1)[:, 0, 0]
    Xs = (Xt - o["mx"]) / o["sx"]
    Tt = Xs @ o["R"]
    ref = oracle.lda(o["T"], lab, 20, 15, Tt)
    d = np.sort(ref["disc"], axis=1)
    margin = d[:, -1] - d[:, -2]
    ok = (got == ref["labels"]) | (margin < 1e-6)
    assert ok.all(), np.flatnonzero(~ok)[:10]
    m.close()


# ----------------------------------------------------------------- PLS-SVD
@pytest.mark.parametrize("name", ["plssvd_sketch", "plssvd_dense", "plssvd_labels"])
def test_plssvd_golden(hip, oracle, name):
    g = load_golden(name)
    kw = dict(components=int(g["ncomp"]), oversample=int(g["oversample"]), power=int(g["power"]),
              seed=int(g["seed"]))
    if "labels" in g:
        m = fit(hip, g["X"], labels=g["labels"], classes=int(g["classes"]), method=1, **kw)
    else:
        m = fit(hip, g["X"], g["Y"], method=1, **kw)
    R = m.export("R").astype(np.float64)
    for k in range(1, int(g["ncomp"]) + 1):
        # B_k = R_k W_k ; compare predictions on the training rows
        pred = m.predict_path(g["X"], [k])[:, :, 0]
        Xs = (g["X"] - g["mx"]) / g["sx"]
        yh = Xs @ (g["R"][:, :k] @ g[f"W{k}"]) + g["my"]
        assert rel(pred, yh) < 1e-8, k
    assert rel(align_columns(R, g["R"]), g["R"]) < 1e-8
    m.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_plssvd_randomized_midsize(hip, oracle, precision):
    """PLS-SVD through the randomized SVD of S (k + oversample < min(p, q))."""
    X, Y = latent_data(3000, 400, 50, factors=30, noise=0.05, seed=5, decay=0.8)
    Xp = X.astype(np.float32 if precision == 32 else np.float64)
    Yp = Y.astype(Xp.dtype)
    m = fit(hip, Xp, Yp, method=1, precision=precision, components=20, oversample=10, power=2,
            seed=1)
    o = oracle.plssvd_fit(Xp.astype(np.float64), Yp.astype(np.float64), ncomp=20, oversample=10,
                          power=2, seed=1)
    pred = m.predict_path(Xp, [5, 20]).astype(np.float64)
    Xs = (Xp.astype(np.float64) - o["mx"]) / o["sx"]
    for j, k in enumerate((5, 20)):
        yh = Xs @ (o["R"][:, :k] @ o["W"][k - 1]) + o["my"]
        assert rel(pred[:, :, j], yh) < TOL[precision] * 10, k
    m.close()


# -------------------------------------------------------------- prediction
def test_predict_classify_project_response_sums(hip, oracle):
    X, lab = class_data(1200, 80, 5, latent=6, seed=9)
    m = fit(hip, X, labels=lab, classes=5, components=6, seed=3)
    o = oracle.simpls_fit(X, labels=lab, classes=5, ncomp=6, seed=3)
    Xt, labt = class_data(300, 80, 5, latent=6, seed=10)
    Xs = (Xt - o["mx"]) / o["sx"]
    # projection
    sc = m.project(Xt, 4)
    assert rel(align_columns(sc, Xs @ o["R"][:, :4]), Xs @ o["R"][:, :4]) < 1e-9
    # regression-style predictions and top-k
    pr = m.predict_path(Xt, [2, 6])
    for j, k in enumerate((2, 6)):
        yh = Xs @ coef(o, k) + o["my"]
        assert rel(pr[:, :, j], yh) < 1e-9
    top = m.classify_path(Xt, [2, 6], lda=False, top=3)
    for j, k in enumerate((2, 6)):
        yh = Xs @ coef(o, k) + o["my"]
        np.testing.assert_array_equal(top[:, :, j], oracle.topk(yh, 3))
    labels, resp = m.classify_response_path(Xt, [6], lda=True, top=2)
    assert rel(resp[:, :, 0], Xs @ coef(o, 6) + o["my"]) < 1e-9
    ref = oracle.lda(o["T"], lab, 5, 6, Xs @ o["R"])
    np.testing.assert_array_equal(labels[:, 0, 0], ref["labels"])
    disc = m.predict_path(Xt, [6], lda=True)[:, :, 0]
    assert rel(disc, ref["disc"]) < 1e-9
    # response sums against the oracle's brute force
    rs = m.response_sums(Xt, 6, labels=labt)
    yh = Xs @ coef(o, 6) + o["my"]
    ref_rs = oracle.response_sums(yh, o["my"], labels=labt)
    np.testing.assert_allclose(rs, ref_rs, rtol=1e-9)
    m.close()


def test_export_variance_controls_compact(hip, oracle):
    X, Y = latent_data(500, 60, 4, factors=5, seed=12)
    m = fit(hip, X, Y, components=5, seed=1)
    o = oracle.simpls_fit(X, Y, ncomp=5, seed=1)
    Xs = (X - o["mx"]) / o["sx"]
    v = oracle.variance(Xs, o["T"])
    P = m.export("P").astype(np.float64)
    ss = m.export("variance").astype(np.float64)
    assert rel(align_columns(P, v["P"]), v["P"]) < 1e-9
    np.testing.assert_allclose(ss, v["ss"], rtol=1e-9)
    ctl = m.controls()
    assert ctl["oversample"] == 32 and ctl["power"] == 5 and ctl["block"] == 1
    m.compact()
    with pytest.raises(hip[0].HipError):
        m.export("T")
    # prediction still works after compaction
    pr = m.predict_path(X, [5])[:, :, 0]
    assert rel(pr, Xs @ coef(o, 5) + o["my"]) < 1e-9
    m.close()


def test_gemm_matches_numpy(hip):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((300, 70)); b = rng.standard_normal((70, 90))
    # FP32 products run in bf16x6 split precision (three bf16 terms per
    # operand, each 32-long reduction step an FP32 chain folded into FP64):
    # the error is that of rounding the inputs and the output to FP32, about
    # 1e-7 relative. (The reference's FP32 GEMM runs in TF32, unit roundoff
    # 2^-11; cuda_resident_rsvd.cuh:20-23.)
    for precision in (64, 32):
        out = hip[0].gemm(a, b, precision).astype(np.float64)
        assert rel(out, a @ b) < (1e-13 if precision == 64 else 1e-6)


# ------------------------------------------------------------------ errors
def test_errors_are_reported_not_fallen_back(hip):
    fp = hip[0]
    X, lab = class_data(100, 10, 3, latent=2, seed=1)
    bad = lab.copy(); bad[5] = 7
    with pytest.raises(fp.HipError, match="labels"):
        fit(hip, X, labels=bad, classes=3, components=2)
    empty = lab.copy(); empty[empty == 3] = 1
    with pytest.raises(fp.HipError, match="class"):
        fit(hip, X, labels=empty, classes=3, components=2)
    Xn = X.copy(); Xn[3, 4] = np.nan
    with pytest.raises(fp.HipError):
        fit(hip, Xn, labels=lab, classes=3, components=2)
    m = fit(hip, X, labels=lab, classes=3, components=2)
    with pytest.raises(fp.HipError, match="increasing"):
        m.predict_path(X, [2, 1])
    with pytest.raises(fp.HipError):
        m.predict_path(X, [3])
    m.close()
    with pytest.raises(fp.HipError, match="dimensions"):
        fit(hip, X, labels=lab, classes=3, components=50)
    with pytest.raises(fp.HipError, match="components must be below"):
        fit(hip, X, labels=lab, classes=3, method=1, components=3)


def test_plsda_full_size(hip):
    """PLS-DA at full size: n=50000, p=10000, K=20 balanced classes whose means
    are N(0, 0.5² I) in a 15-dimensional latent space, X = Z Lᵀ + E, 15
    SIMPLS components and LDA; 10000 held-out rows. FP32 storage against the
    FP64 path on the same values (both held to the oracle at 10000 x 2500 in
    test_simpls_block_routing_classification): coefficients within 1e-5,
    discriminant scores within 1e-5, held-out labels identical except where
    the FP64 top-2 margin is below 1e-6."""
    fp = hip[0]
    n, p, K, A, nt = 50000, 10000, 20, 15, 10000
    rng = np.random.default_rng(2)
    labels = (np.arange(n) % K + 1).astype(np.int32)
    rng.shuffle(labels)
    means = 0.5 * rng.standard_normal((K, 15))
    L = rng.standard_normal((p, 15))

    def draw(lab, r):
        X = (means[lab - 1] @ L.T).astype(np.float32)
        X += r.standard_normal(X.shape, dtype=np.float32)
        return np.asfortranarray(X)

    X = draw(labels, rng)
    rt = np.random.default_rng([2, 2])
    labt = (np.arange(nt) % K + 1).astype(np.int32)
    Xt = draw(labt, rt)
    m32 = fp.ResidentModel.fit(X, labels=labels, classes=K, precision=32, components=A, seed=1)
    m64 = fp.ResidentModel.fit(X.astype(np.float64), labels=labels, classes=K, precision=64,
                               components=A, seed=1)
    assert m32.controls()["block"] == m64.controls()["block"] == 15
    for k in (1, 8, 15):
        assert rel(hip_coef(m32, k), hip_coef(m64, k)) < 1e-5, k
    d32 = m32.predict_path(Xt, [A], lda=True)[:, :, 0].astype(np.float64)
    d64 = m64.predict_path(Xt.astype(np.float64), [A], lda=True)[:, :, 0]
    assert rel(d32, d64) < 1e-5
    l32 = m32.classify_path(Xt, [A], lda=True, top=1)[:, 0, 0]
    l64 = m64.classify_path(Xt.astype(np.float64), [A], lda=True, top=1)[:, 0, 0]
    s = np.sort(d64, axis=1)
    ok = (l32 == l64) | (s[:, -1] - s[:, -2] < 1e-6)
    assert ok.all(), np.flatnonzero(~ok)[:10]
    assert (l64 == labt).mean() > 0.5   # the classes are separable at this spread
    m32.close()
    m64.close()
