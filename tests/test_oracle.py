"""Pinning the FP64 oracle (CPU only).

Each oracle function is checked against something other than itself:
  * golden vectors written by the reference's own header core compiled in
    oracle/_ref (tests/golden/make_golden.py);
  * the reference's known-answer tests, restated (core/tests/*.cpp);
  * closed forms, invariants and brute force.
"""
import numpy as np
import pytest

from conftest import align_columns, load_golden, rel
from oracle.oracle import latent_data


def coef(m, k):
    return m["R"][:, :k] @ m["Q"][:, :k].T


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("name", ["simpls_dense_small", "simpls_dense_sketch", "simpls_labels_small"])
def test_simpls_matches_reference_core(oracle, name):
    g = load_golden(name)
    kw = dict(ncomp=int(g["ncomp"]), oversample=int(g["oversample"]), power=int(g["power"]),
              seed=int(g["seed"]))
    if "labels" in g:
        m = oracle.simpls_fit(g["X"], labels=g["labels"], classes=int(g["classes"]), **kw)
    else:
        m = oracle.simpls_fit(g["X"], g["Y"], **kw)
    assert m["completed"] == int(g["completed"])
    np.testing.assert_allclose(m["mx"], g["mx"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m["sx"], g["sx"], rtol=1e-13)
    np.testing.assert_allclose(m["my"], g["my"], rtol=1e-13, atol=1e-15)
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(coef(m, k), coef(g, k)) < 1e-8, k
    # components agree individually up to sign
    assert rel(align_columns(m["R"], g["R"]), g["R"]) < 1e-7
    assert rel(align_columns(m["T"], g["T"]), g["T"]) < 1e-7


def test_simpls_block_geometry_matches_reference_core(oracle):
    g = load_golden("simpls_block3")
    prep = oracle.prepare(g["X"], labels=g["labels"], classes=int(g["classes"]))
    m = oracle.simpls_core(prep["xs"], prep["S"], int(g["ncomp"]), block=int(g["block"]),
                           batch_geometry=int(g["batch_geometry"]), oversample=int(g["oversample"]),
                           power=int(g["power"]), seed=int(g["seed"]))
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(coef(m, k), coef(g, k)) < 1e-8


def test_simpls_c1_workload_matches_reference_core(oracle):
    g = load_golden("simpls_c1")
    n, p, q, factors, noise, seed = g["recipe"]
    X, Y = latent_data(int(n), int(p), int(q), factors=int(factors), noise=noise, seed=int(seed))
    import hashlib
    assert hashlib.sha256(np.asfortranarray(X).tobytes(order="F")).hexdigest() == str(g["x_sha256"])
    m = oracle.simpls_fit(X, Y, ncomp=int(g["ncomp"]), seed=int(g["seed"]))
    for k in (1, 5, 10):
        assert rel(coef(m, k), coef(g, k)) < 1e-9


def test_simpls_reorth_matches_reference_core(oracle):
    """Dense SIMPLS with the re-orthogonalisation forced (the reference float
    core's routing: scores and weights against the earlier scores twice,
    simpls.hpp:161-200, 635-642, loadings twice) against the compiled
    reference core, at 30 components of a rank-8 problem. The single-pass
    routing parts from the fixture by far more than the tolerance there, so the
    check pins the re-orthogonalisation itself."""
    g = load_golden("simpls_dense_reorth")
    n, p, q, factors, noise, seed, decay = g["recipe"]
    X, Y = latent_data(int(n), int(p), int(q), factors=int(factors), noise=float(noise),
                       seed=int(seed), decay=float(decay))
    import hashlib
    assert hashlib.sha256(np.asfortranarray(X).tobytes(order="F")).hexdigest() == str(g["x_sha256"])
    A = int(g["ncomp"])
    kw = dict(ncomp=A, oversample=int(g["oversample"]), power=int(g["power"]), seed=int(g["seed"]))
    m1 = oracle.simpls_fit(X, Y, reorth=1, **kw)
    m0 = oracle.simpls_fit(X, Y, reorth=0, **kw)
    assert m1["completed"] == int(g["completed"]) == A
    for k in range(1, A + 1):
        assert rel(coef(m1, k), coef(g, k)) < 1e-10, k
    assert rel(coef(m0, A), coef(g, A)) > 1e-8
    np.testing.assert_allclose(m1["T"].T @ m1["T"], np.eye(A), atol=1e-13)


@pytest.mark.parametrize("name", ["plssvd_sketch", "plssvd_dense", "plssvd_labels"])
def test_plssvd_matches_reference_core(oracle, name):
    g = load_golden(name)
    kw = dict(ncomp=int(g["ncomp"]), oversample=int(g["oversample"]), power=int(g["power"]),
              seed=int(g["seed"]))
    if "labels" in g:
        m = oracle.plssvd_fit(g["X"], labels=g["labels"], classes=int(g["classes"]), **kw)
    else:
        m = oracle.plssvd_fit(g["X"], g["Y"], **kw)
    np.testing.assert_allclose(m["sigma"], g["sigma"], rtol=1e-10)
    for k in range(1, int(g["ncomp"]) + 1):
        Bk = m["R"][:, :k] @ m["W"][k - 1]
        Bg = g["R"][:, :k] @ g[f"W{k}"]
        assert rel(Bk, Bg) < 1e-9, k


def test_rsvd_matches_reference_core(oracle):
    g = load_golden("rsvd_small")
    r = oracle.rsvd(g["A"], int(g["k"]), oversample=int(g["oversample"]), power=int(g["power"]),
                    seed=int(g["seed"]))
    np.testing.assert_allclose(r["s"], g["s"], rtol=1e-12)
    assert rel(r["U"] @ r["U"].T, g["U"] @ g["U"].T) < 1e-10
    assert rel(align_columns(r["Vt"].T, g["Vt"].T), g["Vt"].T) < 1e-10


AUDIT_INTS = ("certified", "deterministic_fallback", "attempts", "effective_oversample",
              "effective_power", "effective_seed")


@pytest.mark.parametrize("name", ["audited_lowrank", "audited_flat", "audited_tiny",
                                  "audited_wide", "audited_left"])
def test_audited_rsvd_matches_reference_core(oracle, name):
    """fastsvd: the same attempt/recovery decisions and the same triplets as
    the compiled reference core (audited_rsvd.hpp:298-414)."""
    from oracle.oracle import AUDIT_FIELDS
    g = load_golden(name)
    left = bool(g["left_only"])
    r = oracle.audited_rsvd(g["A"], int(g["k"]), oversample=int(g["oversample"]),
                            power=int(g["power"]), seed=int(g["seed"]), left_only=left)
    gaudit = dict(zip(AUDIT_FIELDS, g["audit"].tolist()))
    for f in AUDIT_INTS:
        assert r["audit"][f] == gaudit[f], f
    np.testing.assert_allclose(r["s"], g["s"], rtol=1e-12)
    assert rel(r["U"] @ r["U"].T, g["U"] @ g["U"].T) < 1e-10
    if not left:
        assert rel(align_columns(r["Vt"].T, g["Vt"].T), g["Vt"].T) < 1e-10
    for f in ("triplet_residual", "omitted_direction_ratio"):
        assert r["audit"][f] == pytest.approx(gaudit[f], rel=1e-6, abs=1e-12), f


def test_audited_rsvd_recovery_slow_spectrum(oracle):
    """tests/testthat/test-rsvd-native-recovery.R: σ = 1 - 0.005 k, no
    oversampling, no power steps: certified only after recovery widenings
    (more than three attempts, oversample above 48, power at least 6, a new
    seed), σ to 2e-5 and triplet residuals at 2e-12."""
    rng = np.random.default_rng(614)
    U0 = np.linalg.qr(rng.standard_normal((160, 100)))[0]
    V0 = np.linalg.qr(rng.standard_normal((125, 100)))[0]
    sv = 1 - 0.005 * np.arange(100)
    A = (U0 * sv) @ V0.T
    for seed in (1, 7):
        r = oracle.audited_rsvd(A, 6, oversample=0, power=0, seed=seed)
        a = r["audit"]
        assert a["certified"] == 1 and a["deterministic_fallback"] == 0
        assert a["attempts"] > 3 and a["effective_oversample"] > 48 and a["effective_power"] >= 6
        assert a["effective_seed"] != seed
        assert a["triplet_residual"] <= 1e-2 and a["omitted_direction_ratio"] <= 1.01
        np.testing.assert_allclose(r["s"], sv[:6], rtol=2e-5)
        V = r["Vt"].T
        res = [max(np.linalg.norm(A @ V[:, j] - r["s"][j] * r["U"][:, j]),
                   np.linalg.norm(A.T @ r["U"][:, j] - r["s"][j] * V[:, j])) / r["s"][0]
               for j in range(6)]
        assert max(res) <= 2e-12


def test_audited_rsvd_closed_forms(oracle):
    """Exact low-rank input: the audited triplets are the top singular
    triplets of the dense SVD; the exact route below six columns equals it."""
    rng = np.random.default_rng(61)
    A = rng.standard_normal((80, 6)) @ rng.standard_normal((6, 50))
    r = oracle.audited_rsvd(A, 4, oversample=4, power=1, seed=2)
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    np.testing.assert_allclose(r["s"], s[:4], rtol=1e-10)
    assert rel(np.abs(r["U"].T @ U[:, :4]), np.eye(4)) < 1e-8
    assert r["audit"]["certified"] == 1 and r["audit"]["attempts"] == 1
    B = rng.standard_normal((5, 30))
    t = oracle.audited_rsvd(B, 5)
    np.testing.assert_allclose(t["s"], np.linalg.svd(B, compute_uv=False), rtol=1e-12)
    assert t["audit"]["deterministic_fallback"] == 1


def test_fastcor_closed_forms(oracle):
    """fastcor (src/r_api.cpp:4054-4142) is the Pearson correlation of units:
    numpy corrcoef for columns and rows, its diagonal, and NaN for a constant
    or non-finite unit."""
    rng = np.random.default_rng(62)
    a = rng.standard_normal((40, 6)); b = rng.standard_normal((40, 8))
    ref = np.corrcoef(a.T, b.T)[:6, 6:]
    np.testing.assert_allclose(oracle.fastcor(a, b), ref, atol=1e-14)
    np.testing.assert_allclose(oracle.fastcor(a.T.copy(), b.T.copy(), by_row=True), ref, atol=1e-14)
    np.testing.assert_allclose(oracle.fastcor(a), np.corrcoef(a.T), atol=1e-14)
    c = rng.standard_normal((40, 6))
    d = oracle.fastcor(a, c, diagonal=True)
    np.testing.assert_allclose(d, [np.corrcoef(a[:, j], c[:, j])[0, 1] for j in range(6)],
                               atol=1e-14)
    bad = a.copy(); bad[:, 2] = 3.0; bad[5, 4] = np.nan
    out = oracle.fastcor(bad, b)
    assert np.all(np.isnan(out[[2, 4]])) and np.all(np.isfinite(out[[0, 1, 3, 5]]))


def test_lda_matches_reference_core(oracle):
    g = load_golden("lda_small")
    out = oracle.lda(g["T"], g["labels"], int(g["classes"]), int(g["prefix"]), g["test"])
    np.testing.assert_allclose(out["disc"], g["disc"], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(out["labels"], g["predicted"])
    assert out["ridge"] == pytest.approx(float(g["ridge"]), rel=1e-14)


# ------------------------------------------------- restated known answers
def _kat_plssvd_data():
    # core/tests/test_core_plssvd.cpp:11-25
    x = np.zeros((8, 3)); y = np.zeros((8, 2))
    for r in range(8):
        v = r - 3.5
        x[r] = [v, v * v - 5.25, -1.0 if r % 2 == 0 else 1.0]
        y[r] = [2 * x[r, 0] - 0.5 * x[r, 1], x[r, 2] + 0.25 * x[r, 0]]
    return x, y


def test_kat_plssvd_sse_and_sigma(oracle):
    # core/tests/test_core_plssvd.cpp:26-91: SSE 10.5762741916562, sigma1 119.561004785789
    x, y = _kat_plssvd_data()
    S = x.T @ y
    m = oracle.plssvd_core(x, S, [1, 2], oversample=4, power=2, seed=17)
    pred = m["T"][:, :2] @ m["W_last"]
    sse = float(((pred - y) ** 2).sum())
    assert abs(sse - 10.5762741916562) < 2e-7
    assert abs(m["sigma"][0] - 119.561004785789) < 2e-7


def test_kat_simpls_exact_recovery(oracle):
    # core/tests/test_core_simpls.cpp:11-117
    vals = [-3, -2, -1, 0, 0, 1, 2, 3, 1, -1, 1, -1, 1, -1, 1, -1, -1, 2, -2, 1, 1, -2, 2, -1]
    x = np.array(vals, dtype=np.float64).reshape((8, 3), order="F")
    coefs = np.zeros((3, 3))
    coefs[0, 0], coefs[1, 0], coefs[0, 1], coefs[1, 1], coefs[0, 2], coefs[1, 2] = 1.5, -0.5, -0.75, 2, 0.25, 1.25
    y = x @ coefs
    S = x.T @ y
    m = oracle.simpls_core(x, S, 2, block=1, reorth=1, oversample=0, power=5, seed=19)
    assert m["completed"] == 2
    pred = m["T"] @ (m["Q"].T)
    # X y-part lies in a 2-dim subspace: relative error < 1e-8 (f64 contract)
    assert np.sum((pred - y) ** 2) / np.sum(y ** 2) < 1e-8
    np.testing.assert_allclose(m["T"].T @ m["T"], np.eye(2), atol=1e-10)
    np.testing.assert_allclose(m["V"].T @ m["V"], np.eye(2), atol=1e-10)


def test_kat_rsvd_diagonal(oracle):
    # core/tests/test_core_operator_rsvd.cpp:141-180: diag(8..1) padded to 10x8
    a = np.zeros((10, 8))
    for i in range(8):
        a[i, i] = 8 - i
    r = oracle.rsvd(a, 2, oversample=6, power=1, seed=27)
    np.testing.assert_allclose(r["s"], [8, 7], atol=1e-8)
    resid = a @ r["Vt"].T - r["U"] * r["s"]
    assert np.abs(resid).max() < 1e-8


# ------------------------------------------------------------ closed forms
def test_preprocessing_closed_form(oracle):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((50, 7)) * 3 + 10
    X[:, 3] = 4.0  # constant column: scale falls back to 1
    Y = rng.standard_normal((50, 2))
    for scaling in (1, 2, 3):
        pr = oracle.prepare(X, Y, scaling=scaling)
        mu = X.mean(0) if scaling != 3 else np.zeros(7)
        sd = X.std(0, ddof=1) if scaling == 2 else np.ones(7)
        sd = np.where(sd > 1e-9, sd, 1.0)
        xs = (X - mu) / sd
        np.testing.assert_allclose(pr["xs"], xs, atol=1e-12)
        np.testing.assert_allclose(pr["S"], xs.T @ (Y - Y.mean(0)), atol=1e-10)


def test_label_preprocessing_equals_centred_one_hot(oracle):
    rng = np.random.default_rng(4)
    X = rng.standard_normal((60, 5))
    lab = rng.integers(1, 5, size=60).astype(np.int32)
    lab[:4] = [1, 2, 3, 4]
    pr = oracle.prepare(X, labels=lab, classes=4)
    H = np.eye(4)[lab - 1]
    xs = (X - X.mean(0)) / X.std(0, ddof=1)
    np.testing.assert_allclose(pr["S"], xs.T @ (H - H.mean(0)), atol=1e-12)
    np.testing.assert_allclose(pr["my"], H.mean(0))


def test_qr_eigen_closed_forms(oracle):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((30, 6))
    q = oracle.qr(a)
    np.testing.assert_allclose(q.T @ q, np.eye(6), atol=1e-13)
    # same column space: projection of a onto q reproduces a
    np.testing.assert_allclose(q @ (q.T @ a), a, atol=1e-12)
    s = a.T @ a
    vals, vecs = oracle.eigen(s)
    np.testing.assert_allclose(vals, np.linalg.eigvalsh(s), rtol=1e-12)
    np.testing.assert_allclose(s @ vecs, vecs * vals, atol=1e-10)


def test_simpls_width_covers_column_space_equals_exact_svd_direction(oracle):
    # q <= oversample+1: the first direction is the top left singular vector of S
    X, Y = latent_data(80, 15, 3, factors=4, seed=6)
    m = oracle.simpls_fit(X, Y, ncomp=1, oversample=32, power=5, seed=9)
    pr = oracle.prepare(X, Y)
    u = np.linalg.svd(pr["S"])[0][:, 0]
    r = m["R"][:, 0] / np.linalg.norm(m["R"][:, 0])
    assert abs(abs(r @ u) - 1.0) < 1e-12


def test_simpls_square_sketch_independent_of_sketch(oracle):
    # every refresh has width q (q <= block + oversample): the sketch spans
    # ℝ^q, so the whole fit is independent of the Gaussian draws and of the
    # power iterations (the premise of the square-sketch refresh); with a
    # narrower sketch it is not (rounding differences reach ~1e-10 in the
    # components past the data's 6 factors)
    X, Y = latent_data(200, 40, 4, factors=6, noise=0.2, seed=11)
    fits = [oracle.simpls_fit(X, Y, ncomp=12, oversample=o, power=pw, seed=sd)
            for o, pw, sd in ((4, 2, 3), (4, 0, 77), (9, 5, 1234))]
    for k in range(1, 13):
        B = [f["R"][:, :k] @ f["Q"][:, :k].T for f in fits]
        assert rel(B[1], B[0]) < 1e-8 and rel(B[2], B[0]) < 1e-8, k
    narrow = [oracle.simpls_fit(X, Y, ncomp=3, oversample=0, power=0, seed=sd) for sd in (3, 4)]
    assert rel(narrow[0]["R"] @ narrow[0]["Q"].T, narrow[1]["R"] @ narrow[1]["Q"].T) > 1e-6


def test_topk_tie_rule(oracle):
    vals = np.array([[1.0, 3.0, 3.0, 2.0], [5.0, 5.0, 5.0, 5.0]])
    out = oracle.topk(vals, 3)
    np.testing.assert_array_equal(out[0], [2, 3, 4])
    np.testing.assert_array_equal(out[1], [1, 2, 3])


def test_response_sums_brute_force(oracle):
    rng = np.random.default_rng(7)
    y = rng.standard_normal((40, 3)); pred = y + 0.1 * rng.standard_normal((40, 3))
    tm = np.array([0.1, -0.2, 0.3])
    out = oracle.response_sums(pred, tm, observed=y)
    np.testing.assert_allclose(out[0], ((y - pred) ** 2).sum(0))
    np.testing.assert_allclose(out[1], ((y - tm) ** 2).sum(0))
    np.testing.assert_allclose(out[2], ((y - y.mean(0)) ** 2).sum(0))


def test_variance_orthogonal_scores_closed_form(oracle):
    # With orthonormal scores the sequential loadings equal the plain ones and
    # the gains are the squared loading norms.
    rng = np.random.default_rng(8)
    xs = rng.standard_normal((50, 6))
    t, _ = np.linalg.qr(rng.standard_normal((50, 3)))
    v = oracle.variance(xs, t)
    np.testing.assert_allclose(v["P"], xs.T @ t, atol=1e-12)
    np.testing.assert_allclose(v["ss"][:3], ((xs.T @ t) ** 2).sum(0), rtol=1e-12)
    assert v["ss"][3] == pytest.approx((xs ** 2).sum())


def test_gaussian_stream_is_the_reference_stream(oracle):
    # first draws of std::mt19937(1) + std::normal_distribution<double>
    a = oracle.gaussian(1, 6)
    b = oracle.gaussian(1, 6)
    np.testing.assert_array_equal(a, b)
    assert abs(a.mean()) < 3 and np.isfinite(a).all()
    g = load_golden("simpls_dense_sketch")
    # the sketch-dependent golden fit only matches with the same stream
    m = oracle.simpls_fit(g["X"], g["Y"], ncomp=int(g["ncomp"]), oversample=int(g["oversample"]),
                          power=int(g["power"]), seed=int(g["seed"]) + 1)
    assert rel(coef(m, 8), coef(g, 8)) > 1e-6


def _reference_or_skip():
    import os
    from oracle.oracle import REF_SO, Reference
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    return Reference()


def test_block_routing_matches_reference_core(oracle):
    # simpls_candidate_block_size (simpls.hpp:43-62) compiled from the header,
    # over a grid that straddles every threshold: work 5e8, responses 1 / 2048,
    # remaining 4, cross-covariance 64 Mi elements.
    ref = _reference_or_skip()
    grid = [(n, p, q, A, cls)
            for n in (10, 2000, 50000, 100000)
            for p in (3, 100, 10000, 20000)
            for q in (1, 2, 3, 20, 2048, 2049, 4000)
            for A in (1, 3, 4, 15, 80)
            for cls in (False, True)]
    for n, p, q, A, cls in grid:
        got = oracle.controls(n, p, q, A, classification=cls)["block"]
        want = ref.block_size(A, p, q, cls, n, 64)
        assert got == want, (n, p, q, A, cls, got, want)
    # the wide-regression branch (p q > 64 Mi) caps the block at 8
    assert ref.block_size(30, 40000, 2000, False, 100, 64) == 8
    assert oracle.controls(100, 40000, 2000, 30)["block"] == 8


def test_crossprod_cache_rule_cases(oracle):
    # src/r_api.cpp:1165-1176 (not compilable here: needs R): the X^T X cache
    # is used iff A >= 20, p <= n, n >= 8 p and p <= 512; batch geometry iff
    # block > 1 and no cache.
    assert oracle.controls(5000, 500, 3, 20)["cache"] == 1
    assert oracle.controls(5000, 500, 3, 19)["cache"] == 0
    assert oracle.controls(3999, 500, 3, 20)["cache"] == 0
    assert oracle.controls(8000, 513, 3, 20)["cache"] == 0
    c = oracle.controls(50000, 10000, 20, 15, classification=True)
    assert (c["block"], c["cache"], c["batch_geometry"]) == (15, 0, 1)
    c = oracle.controls(2000, 1000, 3, 10)
    assert (c["block"], c["cache"], c["batch_geometry"]) == (1, 0, 0)


# ------------------------------------------------------ matrix-free path
@pytest.mark.parametrize("name", ["simpls_operator_block", "simpls_operator_rank1"])
def test_simpls_operator_matches_reference_core(oracle, name):
    """Matrix-free SIMPLS (simpls.hpp:768-1047) against the compiled reference
    core with its centred cross-covariance / projected operators: the block
    route (operator rSVD of the projected operator) and the rank-one
    dominant-direction route (operator_rsvd.hpp:175-202)."""
    g = load_golden(name)
    kw = {k: int(g[k]) for k in ("block", "rank_one") if k in g}
    kw.update({k: int(g[k]) for k in ("batch_geometry", "reorth") if k in g})
    m = oracle.simpls_operator_core(g["X"], g["Y"], ncomp=int(g["ncomp"]),
                                    oversample=int(g["oversample"]), power=int(g["power"]),
                                    seed=int(g["seed"]), **kw)
    assert m["completed"] == int(g["completed"])
    for k in (1, 3, 6):
        assert rel(m["R"][:, :k] @ m["Q"][:, :k].T, g["R"][:, :k] @ g["Q"][:, :k].T) < 1e-10, k
    assert rel(np.abs(m["T"]), np.abs(g["T"])) < 1e-10
    assert rel(np.abs(m["V"]), np.abs(g["V"])) < 1e-10


def test_plssvd_operator_matches_reference_core(oracle):
    g = load_golden("plssvd_operator")
    m = oracle.plssvd_operator_fit(g["X"], g["Y"], ncomp=int(g["ncomp"]),
                                   oversample=int(g["oversample"]), power=int(g["power"]),
                                   seed=int(g["seed"]))
    np.testing.assert_allclose(m["sigma"], g["sigma"], rtol=1e-12)
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(m["R"][:, :k] @ m["W"][k - 1], g["R"][:, :k] @ g[f"W{k}"]) < 1e-10, k


def test_operator_route_equals_explicit_route(oracle):
    """Below the rank-one threshold the matrix-free route computes the same
    subspaces as the explicit one (the operator rSVD of (I - VVᵀ)S with the
    same sketch spans (SSᵀ)^power S Ω as the right-Gram refresh), so
    coefficients agree to rounding; the rank-one route does not (its power
    iteration is not converged at small power) and converges to it as the
    power grows. The 32 MiB routing rule (R/main.R:3454-3463)."""
    X, Y = latent_data(300, 80, 15, factors=8, noise=0.1, seed=8)
    for ncomp, os_, pw in ((6, 4, 2), (8, 10, 1)):
        e = oracle.simpls_fit(X, Y, ncomp=ncomp, oversample=os_, power=pw, seed=3)
        o = oracle.simpls_operator_fit(X, Y, ncomp=ncomp, oversample=os_, power=pw, seed=3)
        assert o["rank_one"] == 0
        assert rel(o["R"] @ o["Q"].T, e["R"] @ e["Q"].T) < 1e-10
        pe = oracle.plssvd_fit(X, Y, ncomp=5, oversample=os_, power=pw, seed=3)
        po = oracle.plssvd_operator_fit(X, Y, ncomp=5, oversample=os_, power=pw, seed=3)
        np.testing.assert_allclose(po["sigma"], pe["sigma"], rtol=1e-10)
        assert rel(po["R"] @ po["W"][4], pe["R"] @ pe["W"][4]) < 1e-10
    # a decaying spectrum, so that the power iteration settles within a few steps
    X, Y = latent_data(300, 80, 15, factors=8, noise=0.1, seed=8, decay=0.5)
    errs = []
    for pw in (1, 4, 12):
        o = oracle.simpls_operator_fit(X, Y, ncomp=4, oversample=4, power=pw, seed=3,
                                       elem=1 << 30)   # forces the rank-one route
        e = oracle.simpls_fit(X, Y, ncomp=4, oversample=4, power=pw, seed=3)
        assert o["rank_one"] == 1
        errs.append(rel(o["R"] @ o["Q"].T, e["R"] @ e["Q"].T))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-3, errs
    assert not oracle.implicit_route(20000, 209) and oracle.implicit_route(20000, 210)


def test_blas_bound_oracle_matches_plain_loops(oracle):
    """The CPU baseline times the oracle with its products on numpy's bundled
    OpenBLAS dgemm (the reference CPU adapter's BLAS). Bound that way it must
    give the plain-loop oracle's fit (only the summation order differs):
    SIMPLS within the latent rank and PLS-SVD, relative 1e-9."""
    from oracle.oracle import latent_data
    X, Y = latent_data(700, 150, 3, factors=6, noise=0.1, seed=12)
    plain_s = oracle.simpls_fit(X, Y, ncomp=6, oversample=10, power=2, seed=3)
    plain_p = oracle.plssvd_fit(X, Y, ncomp=3, oversample=10, power=2, seed=3)
    try:
        desc = oracle.use_blas(2)
        assert "OpenBLAS" in desc
        blas_s = oracle.simpls_fit(X, Y, ncomp=6, oversample=10, power=2, seed=3)
        blas_p = oracle.plssvd_fit(X, Y, ncomp=3, oversample=10, power=2, seed=3)
    finally:
        oracle.use_blas(None)
    for k in range(1, 7):
        b0 = plain_s["R"][:, :k] @ plain_s["Q"][:, :k].T
        b1 = blas_s["R"][:, :k] @ blas_s["Q"][:, :k].T
        assert np.linalg.norm(b1 - b0) / np.linalg.norm(b0) < 1e-9, k
    for key in ("R", "Q", "T"):
        a, b = plain_p[key], blas_p[key]
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-9, key


# ------------------------------------------------- scaling 1 / 3, edge rules
@pytest.mark.parametrize("name", ["simpls_scaling1", "simpls_scaling3", "simpls_labels_scaling1"])
def test_oracle_simpls_scaling_goldens(oracle, name):
    """Scaling 1 (centring, R's default) and 3 (none) against the compiled
    reference core (supervised.hpp:63-124, classification.hpp:86-184)."""
    g = load_golden(name)
    kw = dict(ncomp=int(g["ncomp"]), scaling=int(g["scaling"]), oversample=int(g["oversample"]),
              power=int(g["power"]), seed=int(g["seed"]))
    if "labels" in g:
        o = oracle.simpls_fit(g["X"], labels=g["labels"], classes=int(g["classes"]), **kw)
    else:
        o = oracle.simpls_fit(g["X"], g["Y"], **kw)
    for k in range(1, int(g["ncomp"]) + 1):
        B0 = g["R"][:, :k] @ g["Q"][:, :k].T
        assert rel(o["R"][:, :k] @ o["Q"][:, :k].T, B0) < 1e-10, k
    np.testing.assert_allclose(o["mx"], g["mx"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(o["sx"], g["sx"], rtol=1e-12)
    if int(g["scaling"]) != 2:
        assert np.all(g["sx"] == 1.0)
    if int(g["scaling"]) == 3:
        assert np.all(g["mx"] == 0.0)


@pytest.mark.parametrize("name", ["plssvd_scaling1", "plssvd_scaling3"])
def test_oracle_plssvd_scaling_goldens(oracle, name):
    g = load_golden(name)
    o = oracle.plssvd_fit(g["X"], g["Y"], ncomp=int(g["ncomp"]), scaling=int(g["scaling"]),
                          oversample=int(g["oversample"]), power=int(g["power"]),
                          seed=int(g["seed"]))
    assert rel(align_columns(o["R"], g["R"]), g["R"]) < 1e-10
    np.testing.assert_allclose(o["sigma"], g["sigma"], rtol=1e-12)
    for k in range(1, int(g["ncomp"]) + 1):
        # B_k = R_k W_k does not depend on the components' signs
        assert rel(o["R"][:, :k] @ o["W"][k - 1], g["R"][:, :k] @ g[f"W{k}"]) < 1e-10, k
    np.testing.assert_allclose(o["mx"], g["mx"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(o["sx"], g["sx"], rtol=1e-12)


def test_oracle_lda_ladder_golden(oracle):
    """The LDA ridge ladder past 1e-8 and its exhaustion (lda.hpp:86-131),
    from moments whose pooled covariance has a negative eigenvalue of
    -6e-6 (the 1e-5 rung) or -0.5 (every rung fails)."""
    g = load_golden("lda_ladder")
    for k in (1, 2, 3):
        o = oracle.lda_moments(g["gram"], g["sums"], g["counts"], int(g["n"]), k)
        assert o["relative"] == float(g[f"relative{k}"]), k
        np.testing.assert_allclose(o["linear"], g[f"linear{k}"], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(o["constants"], g[f"constants{k}"], rtol=1e-12, atol=1e-12)
    assert float(g["relative3"]) == 1e-5
    with pytest.raises(RuntimeError, match="every ridge"):
        oracle.lda_moments(g["gram_fail"], g["sums"], g["counts"], int(g["n"]), 3)


def test_oracle_plssvd_pivoted_golden(oracle):
    """The PLS-SVD prefix solve: Cholesky for prefix 1, the pivoted solve for
    the indefinite prefixes 2-4 (linalg.hpp:104-177), failure for a singular
    Gram."""
    g = load_golden("plssvd_pivoted")
    o = oracle.plssvd_weights(g["G"], g["sigma"], g["Q"])
    np.testing.assert_array_equal(o["route"], g["route"])
    assert list(g["route"]) == [0, 1, 1, 1]
    for k in range(1, 5):
        np.testing.assert_allclose(o["W"][k - 1], g[f"W{k}"], rtol=1e-12, atol=1e-12)
    with pytest.raises(RuntimeError, match="latent solve failed"):
        oracle.plssvd_weights(g["G_singular"], g["sigma"][:2], g["Q"][:, :2])


# ------------------------------------------------------- OPLS and kernel PLS
def opls_predict(m, x, prefix):
    """Ŷ of an oracle OPLS fit by filtering the standardised rows explicitly
    (opls.hpp:467-509) and applying the inner SIMPLS model."""
    xs = (x - m["mx"]) / m["sx"]
    for c in range(m["W"].shape[1]):
        s = xs @ m["W"][:, c]
        xs = xs - np.outer(s, m["P"][:, c])
    return xs @ m["R"][:, :prefix] @ m["Q"][:, :prefix].T + m["my"]


@pytest.mark.parametrize("name", ["opls_dense", "opls_sketch", "opls_labels"])
def test_opls_matches_reference_core(oracle, name):
    g = load_golden(name)
    kw = dict(ncomp=int(g["ncomp"]), north=int(g["north"]), scaling=int(g["scaling"]),
              oversample=int(g["oversample"]), power=int(g["power"]), seed=int(g["seed"]))
    if "labels" in g:
        m = oracle.opls_fit(g["X"], labels=g["labels"], classes=int(g["classes"]), **kw)
    else:
        m = oracle.opls_fit(g["X"], g["Y"], **kw)
    assert m["filtered"] == int(g["filtered"]) and m["completed"] == int(g["completed"])
    np.testing.assert_allclose(m["mx"], g["mx"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m["sx"], g["sx"], rtol=1e-13)
    np.testing.assert_allclose(m["my"], g["my"], rtol=1e-13, atol=1e-15)
    # the filter's weights and loadings, up to a common sign per component
    assert rel(align_columns(m["W"], g["W"]), g["W"]) < 1e-10
    assert rel(align_columns(m["P"], g["P"]), g["P"]) < 1e-10
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(coef(m, k), coef(g, k)) < 1e-9, k
    if "pred" in g:
        k = int(g["prefix"])
        assert rel(opls_predict(m, g["Xt"], k), g["pred"]) < 1e-10
        # R_eff takes standardised rows straight to the inner scores
        xs = (g["Xt"] - m["mx"]) / m["sx"]
        pe = xs @ m["Reff"][:, :k] @ m["Q"][:, :k].T + m["my"]
        assert rel(pe, g["pred"]) < 1e-10


def test_opls_filter_invariants(oracle):
    """The filtered predictors annihilate every orthogonal weight
    (X_f w_o = X w_o (1 - p_oᵀ w_o) with p_oᵀ w_o = 1), and with no filter
    component OPLS is SIMPLS on the same standardised predictors."""
    X, Y = latent_data(120, 25, 3, factors=5, seed=21)
    m = oracle.opls_fit(X, Y, ncomp=3, north=2, oversample=4, power=2, seed=5)
    prep = oracle.prepare(X, Y)
    xf = prep["xs"].copy()
    for c in range(2):
        wo, po = m["W"][:, c], m["P"][:, c]
        assert abs(po @ wo - 1.0) < 1e-12
        xf = xf - np.outer(xf @ wo, po)
    for c in range(2):
        assert np.abs(xf @ m["W"][:, c]).max() < 1e-10 * np.abs(xf).max()
    m0 = oracle.opls_fit(X, Y, ncomp=3, north=0, oversample=4, power=2, seed=5)
    s = oracle.simpls_fit(X, Y, ncomp=3, oversample=4, power=2, seed=5)
    for k in (1, 3):
        assert rel(coef(m0, k), coef(s, k)) < 1e-10
    np.testing.assert_allclose(m0["Reff"], m0["R"], rtol=0, atol=0)


@pytest.mark.parametrize("name", ["kernelpls_rbf", "kernelpls_poly"])
def test_kernelpls_matches_reference_core(oracle, name):
    g = load_golden(name)
    m = oracle.kernelpls_fit(g["X"], g["Y"], ncomp=int(g["ncomp"]), scaling=int(g["scaling"]),
                             kernel=int(g["kernel"]), gamma=float(g["gamma"]),
                             degree=int(g["degree"]), coef0=float(g["coef0"]),
                             oversample=int(g["oversample"]), power=int(g["power"]),
                             seed=int(g["seed"]), xtest=g["Xt"], prefix=int(g["prefix"]))
    assert m["completed"] == int(g["completed"])
    np.testing.assert_allclose(m["mx"], g["mx"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m["sx"], g["sx"], rtol=1e-13)
    np.testing.assert_allclose(m["colmeans"], g["colmeans"], rtol=1e-12, atol=1e-14)
    assert abs(m["grand"] - float(g["grand"])) < 1e-12 * max(1.0, abs(float(g["grand"])))
    for k in range(1, int(g["ncomp"]) + 1):
        assert rel(coef(m, k), coef(g, k)) < 1e-9, k
    assert rel(m["pred"], g["pred"]) < 1e-10


def test_kernelpls_closed_forms(oracle):
    """The centred training kernel has zero row and column sums, so the
    training scores T = K_c R have zero column means; predicting the
    training rows through the test-kernel path (kernel against the training
    rows, test centring) reproduces the fitted values T_k Q_kᵀ + μ_y."""
    X, Y = latent_data(60, 8, 2, factors=4, seed=31)
    for kernel, gamma, degree in ((2, 0.1, 3), (3, 0.2, 2)):
        m = oracle.kernelpls_fit(X, Y, ncomp=3, kernel=kernel, gamma=gamma, degree=degree,
                                 oversample=4, power=2, seed=3, xtest=X, prefix=3)
        assert np.abs(m["T"].mean(axis=0)).max() < 1e-12 * np.abs(m["T"]).max()
        fitted = m["T"] @ m["Q"].T + m["my"]
        assert rel(m["pred"], fitted) < 1e-12
