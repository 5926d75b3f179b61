"""Cross-validation entries (GPU, through the C ABI) against fold-by-fold
restatements on the FP64 oracle (src/cuda_resident_api.cuh:538-1023).

What the restatements fix, fold by fold: held-out rows are the rows carrying
the fold id, training rows the rest in row order; SIMPLS folds share the
seed, PLS-SVD fold f (0-based) uses seed + f; classification folds drop the
classes absent from their training rows and renumber the rest 1..K' (a fold
whose training rows hold one class predicts it, status 4; a fold with no
rows is status 2); top-1 is the first maximum of the response scores or of
the LDA discriminants; scores land in the columns of the original classes
(zeros elsewhere); regression metrics are SSE over held-out rows against
the training-mean SS (Q²), the total SS of all rows (observed R²) and the
held-out value count (RMSD).

Tolerances: precision 64, 1e-9 relative on metrics, predictions and scores
and identical labels; precision 32 (stored FP32 inputs, FP64 small math),
1e-5 on response scores against the oracle on the same FP32 values and
identical labels except rows whose top-2 margin is below 1e-4.
"""
import numpy as np
import pytest

from conftest import rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu


def _oracle_cv_regression(oracle, method, X, Y, folds, prefixes, seed, oversample, power,
                          scaling=2):
    n, q = Y.shape
    F = folds.max()
    sse = np.zeros(len(prefixes)); tss = np.zeros(len(prefixes)); cnt = np.zeros(len(prefixes))
    pred = np.zeros((n, q, len(prefixes)))
    status = np.zeros(F, dtype=np.int32)
    for f in range(F):
        te = np.where(folds == f + 1)[0]
        if te.size == 0:
            status[f] = 2
            continue
        tr = np.where(folds != f + 1)[0]
        fit = oracle.simpls_fit if method == 0 else oracle.plssvd_fit
        m = fit(X[tr], Y[tr], ncomp=max(prefixes), scaling=scaling, oversample=oversample,
                power=power, seed=seed + (f if method == 1 else 0))
        S = (X[te] - m["mx"]) / m["sx"] @ m["R"]
        for j, k in enumerate(prefixes):
            W = m["Q"][:, :k].T if method == 0 else m["W"][k - 1]
            yh = S[:, :k] @ W + m["my"]
            pred[te, :, j] = yh
            sse[j] += ((Y[te] - yh) ** 2).sum()
            tss[j] += ((Y[te] - m["my"]) ** 2).sum()
            cnt[j] += te.size * q
        status[f] = 1
    total = ((Y - Y.mean(0)) ** 2).sum()
    return dict(status=status, sse=sse, q2=1 - sse / tss, rmsd=np.sqrt(sse / cnt),
                observed_r2=1 - sse / total, predictions=pred)


def _oracle_cv_classification(oracle, method, X, lab, folds, K, prefixes, lda, seed, oversample,
                              power, scaling=2):
    n = len(lab)
    F = folds.max()
    P = len(prefixes)
    preds = np.zeros((n, P), dtype=np.int32)
    scores = np.zeros((n, K, P))
    margin = np.full((n, P), np.inf)
    correct = np.zeros(P); cnt = np.zeros(P)
    status = np.zeros(F, dtype=np.int32)
    for f in range(F):
        te = np.where(folds == f + 1)[0]
        if te.size == 0:
            status[f] = 2
            continue
        tr = np.where(folds != f + 1)[0]
        active = np.unique(lab[tr])
        if active.size <= 1:
            fb = int(active[0]) if active.size else 1
            preds[te, :] = fb
            correct += (lab[te] == fb).sum(); cnt += te.size
            status[f] = 4
            continue
        remap = np.zeros(K + 1, dtype=np.int32)
        remap[active] = np.arange(1, active.size + 1)
        comp = remap[lab[tr]]
        fit = oracle.simpls_fit if method == 0 else oracle.plssvd_fit
        m = fit(X[tr], labels=comp, classes=active.size, ncomp=max(prefixes), scaling=scaling,
                oversample=oversample, power=power, seed=seed + (f if method == 1 else 0))
        S = (X[te] - m["mx"]) / m["sx"] @ m["R"]
        for j, k in enumerate(prefixes):
            W = m["Q"][:, :k].T if method == 0 else m["W"][k - 1]
            yh = S[:, :k] @ W + m["my"]
            scores[np.ix_(te, active - 1, [j])] = yh[:, :, None]
            if lda:
                r = oracle.lda(m["T"], comp, active.size, k, S)
                idx, decision = r["labels"], r["disc"]
            else:
                idx, decision = oracle.topk(yh, 1)[:, 0], yh
            top2 = np.sort(decision, axis=1)[:, -2:]
            margin[te, j] = top2[:, 1] - top2[:, 0]
            p = active[idx - 1]
            preds[te, j] = p
            correct[j] += (p == lab[te]).sum(); cnt[j] += te.size
        status[f] = 1
    return dict(status=status, metrics=correct / cnt, predictions=preds, scores=scores,
                margin=margin)


# -------------------------------------------------------------- regression
@pytest.mark.parametrize("metric", [2, 3, 4])
def test_simpls_cv_regression_matches_oracle(hip, oracle, metric):
    fp = hip[0]
    X, Y = latent_data(300, 40, 2, factors=5, noise=0.3, seed=6)
    folds = (np.arange(300) % 5 + 1).astype(np.int32)
    np.random.default_rng(2).shuffle(folds)
    prefixes = [1, 3, 5]
    got = fp.simpls_cv_regression(X, Y, folds, prefixes, precision=64, metric=metric,
                                  oversample=6, power=2, seed=5)
    want = _oracle_cv_regression(oracle, 0, X, Y, folds, prefixes, seed=5, oversample=6, power=2)
    np.testing.assert_array_equal(got["status"], want["status"])
    for key in ("q2", "rmsd", "observed_r2"):
        np.testing.assert_allclose(got[key], want[key], rtol=1e-9)
    expect = {2: want["observed_r2"], 3: want["q2"], 4: want["rmsd"]}[metric]
    np.testing.assert_allclose(got["metrics"], expect, rtol=1e-9)
    assert rel(got["predictions"], want["predictions"]) < 1e-9


@pytest.mark.parametrize("precision,scaling", [(64, 1), (64, 2), (64, 3), (32, 2)])
def test_simpls_cv_regression_row_views(hip, oracle, precision, scaling):
    """The batched regression CV runs each fold as a row view of the one
    resident X: every pass covers all rows with zero responses and zero
    scores on the fold's held-out rows, and the fold's column statistics are
    the full sums minus the held-out rows' (same shift). Uneven folds,
    column means far from zero (the subtraction of shifted sums), ten
    responses (the statistics pass on the FP64 matrix cores); against
    per-fold oracle fits on the gathered training rows (on the FP32 values
    widened to FP64 at precision 32)."""
    fp = hip[0]
    rng = np.random.default_rng(12)
    X, Y = latent_data(700, 300, 10, factors=6, noise=0.2, seed=12)
    X = X + 40.0 * rng.standard_normal(300)
    folds = rng.choice([1, 2, 3, 4], size=700, p=[0.1, 0.2, 0.3, 0.4]).astype(np.int32)
    dt = np.float64 if precision == 64 else np.float32
    Xp, Yp = X.astype(dt), Y.astype(dt)
    prefixes = [1, 4, 8]
    got = fp.simpls_cv_regression(Xp, Yp, folds, prefixes, precision=precision, scaling=scaling,
                                  metric=3, oversample=6, power=2, seed=5)
    want = _oracle_cv_regression(oracle, 0, Xp.astype(np.float64), Yp.astype(np.float64), folds,
                                 prefixes, seed=5, oversample=6, power=2, scaling=scaling)
    np.testing.assert_array_equal(got["status"], want["status"])
    err = rel(got["predictions"], want["predictions"])
    print(f"precision {precision} scaling {scaling}: predictions {err:.2e}")
    tol = 1e-9 if precision == 64 else 1e-7
    assert err < tol
    for key in ("q2", "rmsd", "observed_r2"):
        np.testing.assert_allclose(got[key], want[key], rtol=tol)


def test_simpls_cv_regression_empty_fold_status(hip):
    fp = hip[0]
    X, Y = latent_data(120, 15, 1, factors=3, seed=7)
    folds = np.where(np.arange(120) % 2 == 0, 1, 3).astype(np.int32)  # fold 2 empty
    got = fp.simpls_cv_regression(X, Y, folds, [2], precision=64, metric=3, oversample=4,
                                  power=1, seed=1)
    assert list(got["status"]) == [1, 2, 1]
    with pytest.raises(fp.HipError, match="strictly increasing"):
        fp.simpls_cv_regression(X, Y, folds, [2, 2], precision=64)


def test_plssvd_cv_regression_matches_oracle(hip, oracle):
    fp = hip[0]
    X, Y = latent_data(260, 30, 4, factors=4, noise=0.3, seed=8)
    folds = (np.arange(260) % 4 + 1).astype(np.int32)
    np.random.default_rng(3).shuffle(folds)
    folds[folds == 3] = 5  # fold 3 empty: status 2, seeds still advance by fold index
    prefixes = [1, 2, 4]
    got = fp.plssvd_cv_regression(X, Y, folds, prefixes, precision=64, metric=3, oversample=3,
                                  power=2, seed=11)
    want = _oracle_cv_regression(oracle, 1, X, Y, folds, prefixes, seed=11, oversample=3, power=2)
    np.testing.assert_array_equal(got["status"], want["status"])
    assert list(got["status"]) == [1, 1, 2, 1, 1]
    for key in ("q2", "rmsd", "observed_r2"):
        np.testing.assert_allclose(got[key], want[key], rtol=1e-9)
    np.testing.assert_allclose(got["metrics"], want["q2"], rtol=1e-9)
    assert rel(got["predictions"], want["predictions"]) < 1e-9
    # PLS-SVD components are bounded by the response count
    with pytest.raises(fp.HipError, match="PLS-SVD dimensions"):
        fp.plssvd_cv_regression(X, Y, folds, [5], precision=64)


# ---------------------------------------------------------- classification
def _class_case():
    X, lab = class_data(500, 60, 5, latent=6, seed=21, spread=0.8)
    folds = (np.arange(500) % 4 + 1).astype(np.int32)
    np.random.default_rng(4).shuffle(folds)
    # class 5 only in fold 2: fold 2's training rows lack it, so its fit
    # sees four classes
    lab = lab.copy()
    lab[(lab == 5) & (folds != 2)] = 4
    return X, lab, folds


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("lda", [False, True])
def test_cv_classification_matches_oracle(hip, oracle, method, lda):
    fp = hip[0]
    X, lab, folds = _class_case()
    prefixes = [1, 3] if method == 1 else [1, 3, 6]
    run = fp.simpls_cv_classification if method == 0 else fp.plssvd_cv_classification
    got = run(X, lab, folds, 5, prefixes, precision=64, lda=lda, oversample=4, power=2, seed=9,
              store_scores=True)
    want = _oracle_cv_classification(oracle, method, X, lab, folds, 5, prefixes, lda, seed=9,
                                     oversample=4, power=2)
    np.testing.assert_array_equal(got["status"], want["status"])
    np.testing.assert_array_equal(got["predictions"], want["predictions"])
    np.testing.assert_allclose(got["metrics"], want["metrics"], rtol=1e-12)
    assert rel(got["scores"], want["scores"]) < 1e-9
    # fold 2's fit dropped class 5: that score column is zero on its rows
    assert not np.any(got["scores"][folds == 2, 4, :])
    assert np.all(got["predictions"][folds == 2] != 5)


def test_cv_classification_single_class_and_empty_folds(hip, oracle):
    fp = hip[0]
    X, _ = class_data(60, 12, 2, latent=2, seed=22)
    folds = np.repeat(np.array([1, 2, 4], dtype=np.int32), 20)
    # fold 1 rows are class 2, the rest class 1: fold 1 trains on class 1
    # only (status 4, predicts 1); folds 2 and 4 train on both classes; fold
    # 3 has no rows (status 2)
    lab = np.where(folds == 1, 2, 1).astype(np.int32)
    got = fp.simpls_cv_classification(X, lab, folds, 2, [1], precision=64, oversample=2,
                                      power=1, seed=3)
    want = _oracle_cv_classification(oracle, 0, X, lab, folds, 2, [1], False, seed=3,
                                     oversample=2, power=1)
    assert list(got["status"]) == [4, 1, 2, 1]
    np.testing.assert_array_equal(got["status"], want["status"])
    np.testing.assert_array_equal(got["predictions"], want["predictions"])
    np.testing.assert_allclose(got["metrics"], want["metrics"], rtol=1e-12)
    assert np.all(got["predictions"][folds == 1, 0] == 1)


def test_cv_classification_precision32(hip, oracle):
    fp = hip[0]
    X, lab, folds = _class_case()
    X32 = X.astype(np.float32)
    prefixes = [2, 6]
    got = fp.simpls_cv_classification(X32, lab, folds, 5, prefixes, precision=32, oversample=4,
                                      power=2, seed=9, store_scores=True)
    want = _oracle_cv_classification(oracle, 0, X32.astype(np.float64), lab, folds, 5, prefixes,
                                     False, seed=9, oversample=4, power=2)
    np.testing.assert_array_equal(got["status"], want["status"])
    assert got["scores"].dtype == np.float32
    assert rel(got["scores"], want["scores"]) < 1e-5
    firm = want["margin"] > 1e-4
    np.testing.assert_array_equal(got["predictions"][firm], want["predictions"][firm])


def test_cv_classification_rejects_bad_input(hip):
    fp = hip[0]
    X, lab, folds = _class_case()
    bad = lab.copy(); bad[0] = 6
    with pytest.raises(fp.HipError, match="label or fold"):
        fp.simpls_cv_classification(X, bad, folds, 5, [2], precision=64)
    with pytest.raises(fp.HipError, match="strictly increasing"):
        fp.plssvd_cv_classification(X, lab, folds, 5, [2, 1], precision=64)
    with pytest.raises(fp.HipError, match="below class count"):
        fp.plssvd_cv_classification(X, lab, folds, 5, [4], precision=64)
