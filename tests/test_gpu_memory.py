"""Memory-aware execution (GPU, through the C ABI).

* Queries on host rows go through the device in row blocks once their
  workspace would exceed 2/5 of the free device memory (the reference's
  classify rule, src/cuda_resident_api.cuh:259-297; here predictions too).
  FASTPLS_HIP_QUERY_FREE_BYTES stands for the free memory, so a small value
  forces 1024-row blocks. Blocked predict_path, classify_path and
  classify_response_path results equal the unblocked ones: predictions to
  1e-12 (precision 64) / 1e-6 (precision 32: FP32 outputs) relative (the split
  of the product over p may differ with the block's row count), labels
  exactly (no near-ties in these data).
* The batched SIMPLS regression CV fits its folds in sub-batches that hold at
  most 2/5 of the free memory (FASTPLS_HIP_CV_FREE_BYTES stands for it): with
  one fold per sub-batch its statuses, metrics and predictions match the
  fold-by-fold oracle as the all-folds batch does (1e-9), and the all-folds
  batch to 1e-12.
"""
import os

import numpy as np
import pytest

from conftest import rel
from oracle.oracle import class_data, latent_data
from test_gpu_cv import _oracle_cv_regression

pytestmark = pytest.mark.gpu


class _env:
    def __init__(self, key, value):
        self.key, self.value = key, value

    def __enter__(self):
        os.environ[self.key] = self.value

    def __exit__(self, *exc):
        del os.environ[self.key]


@pytest.mark.parametrize("precision", [64, 32])
def test_blocked_predict_equals_unblocked(hip, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    X, Y = latent_data(900, 120, 3, factors=5, noise=0.2, seed=61)
    Xq, _ = latent_data(3000, 120, 3, factors=5, noise=0.2, seed=62)
    m = fp.ResidentModel.fit(X.astype(dt), Y.astype(dt), precision=precision, components=5,
                             oversample=4, power=1, seed=3)
    whole = m.predict_path(Xq.astype(dt), [1, 3, 5])
    with _env("FASTPLS_HIP_QUERY_FREE_BYTES", "1000000"):
        blocked = m.predict_path(Xq.astype(dt), [1, 3, 5])
    tol = 1e-12 if precision == 64 else 1e-6
    for j in range(3):
        assert rel(blocked[:, :, j], whole[:, :, j]) < tol, j
    m.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_blocked_classify_equals_unblocked(hip, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    X, lab = class_data(1200, 80, 6, latent=5, seed=63, spread=2.0)
    Xq, _ = class_data(2600, 80, 6, latent=5, seed=64, spread=2.0)
    m = fp.ResidentModel.fit(X.astype(dt), labels=lab, classes=6, precision=precision,
                             components=4, oversample=4, power=1, seed=2)
    for lda in (False, True):
        whole = m.classify_path(Xq.astype(dt), [2, 4], lda=lda, top=2)
        wl, wr = m.classify_response_path(Xq.astype(dt), [2, 4], lda=lda, top=3)
        with _env("FASTPLS_HIP_QUERY_FREE_BYTES", "500000"):
            blocked = m.classify_path(Xq.astype(dt), [2, 4], lda=lda, top=2)
            bl, br = m.classify_response_path(Xq.astype(dt), [2, 4], lda=lda, top=3)
        np.testing.assert_array_equal(blocked, whole)
        np.testing.assert_array_equal(bl, wl)
        assert rel(br, wr) < (1e-12 if precision == 64 else 1e-6)
    m.close()


def test_cv_sub_batches_match_oracle(hip, oracle):
    fp = hip[0]
    X, Y = latent_data(300, 40, 2, factors=5, noise=0.3, seed=6)
    folds = (np.arange(300) % 5 + 1).astype(np.int32)
    np.random.default_rng(2).shuffle(folds)
    folds[folds == 4] = 2  # fold 4 empty: status 2
    prefixes = [1, 3, 5]
    kw = dict(precision=64, metric=3, oversample=6, power=2, seed=5)
    whole = fp.simpls_cv_regression(X, Y, folds, prefixes, **kw)
    with _env("FASTPLS_HIP_CV_FREE_BYTES", "1"):
        one = fp.simpls_cv_regression(X, Y, folds, prefixes, **kw)
    want = _oracle_cv_regression(oracle, 0, X, Y, folds, prefixes, seed=5, oversample=6, power=2)
    assert list(one["status"]) == [1, 1, 1, 2, 1]
    np.testing.assert_array_equal(one["status"], want["status"])
    for key in ("q2", "rmsd", "observed_r2", "metrics"):
        np.testing.assert_allclose(one[key], want[key if key != "metrics" else "q2"], rtol=1e-9)
        np.testing.assert_allclose(one[key], whole[key], rtol=1e-12)
    assert rel(one["predictions"], want["predictions"]) < 1e-9
    assert rel(one["predictions"], whole["predictions"]) < 1e-12
