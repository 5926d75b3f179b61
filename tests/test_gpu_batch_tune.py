"""Autotuned split lengths of batched execution (GPU, through the C ABI).

fastpls_hip_batch_autotune re-plans and runs a batch over candidate split
lengths and keeps the fastest plan. Checked, at precision 64 and 32:

* the report holds lengths within 128..4096, at least one plan per tried
  length, and a chosen device time no longer than the starting plan's;
* the batch then reports the chosen lengths, new batches take them as the
  process default, and fastpls_hip_batch_set_splits restores a default;
* the tuned batch holds the same fits: every fit's fitted values T_k Q_kᵀ and
  coefficients B = R_k Q_kᵀ at every prefix agree with the untuned batch's
  to 1e-9 (precision 64) / 1e-6 (precision 32, FP64 products on FP32-stored
  data: only the split order of FP64 sums differs);
* with FASTPLS_HIP_TUNE_FILE set, the choice is written to that file;
* invalid trial counts and lengths are rejected.
"""
import os

import numpy as np
import pytest

from conftest import rel
from oracle.oracle import latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-6}


def _dev(fp, a):
    return fp.DeviceBuffer.from_array(np.asfortranarray(a))


@pytest.mark.parametrize("precision", [64, 32])
def test_autotune_keeps_fits_and_sets_default(hip, precision, tmp_path):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    rng = np.random.default_rng(41)
    builtin = fp.batch_splits()
    keep, groups, fits = [], [], []
    # shared-X permutation refits (matrix-core passes) and VALU singletons
    X, Y = latent_data(3000, 700, 2, factors=6, noise=0.1, seed=42)
    xb = _dev(fp, X.astype(dt)); yb = _dev(fp, Y.astype(dt)); keep += [xb, yb]
    groups.append((xb.ptr, 3000, 700))
    for r in range(10):
        pb = _dev(fp, rng.permutation(3000).astype(np.int32)); keep.append(pb)
        fits.append(dict(group=0, q=2, components=4, seed=60 + r, y=yb.ptr, perm=pb.ptr))
    for i, (n, p) in enumerate([(2500, 1800), (900, 3000), (4000, 400)]):
        Xs, Ys = latent_data(n, p, 3, factors=5, noise=0.2, seed=43 + i)
        xs = _dev(fp, Xs.astype(dt)); ys = _dev(fp, Ys.astype(dt)); keep += [xs, ys]
        groups.append((xs.ptr, n, p))
        fits.append(dict(group=len(groups) - 1, q=3, components=5, seed=70 + i, y=ys.ptr))
    kw = dict(scaling=2, oversample=4, power=1)

    batch = fp.SimplsBatch(precision, groups, fits, **kw)
    before = [{k: batch.export(i, k, f64=True) for k in ("R", "Q", "T")} for i in range(len(fits))]
    path = tmp_path / "splits.txt"
    os.environ["FASTPLS_HIP_TUNE_FILE"] = str(path)
    try:
        rep = batch.autotune(trials=1)
    finally:
        del os.environ["FASTPLS_HIP_TUNE_FILE"]
    assert all(128 <= v <= 4096 for v in rep["splits"]), rep
    assert rep["plans"] >= 2 and rep["seconds"] <= rep["start_seconds"] * (1 + 1e-12), rep
    assert batch.splits() == rep["splits"]
    assert fp.batch_splits() == rep["splits"]
    words = path.read_text().split()
    assert words[0] == "fastpls-hip-batch-splits-v1"
    assert tuple(int(w) for w in words[1:5]) == rep["splits"]
    done, flags, _ = batch.status()
    assert all(f == 0 for f in flags) and done == [f["components"] for f in fits]
    for i, f in enumerate(fits):
        R, Q, T = (batch.export(i, k, f64=True) for k in ("R", "Q", "T"))
        b = before[i]
        for k in range(1, f["components"] + 1):
            assert rel(T[:, :k] @ Q[:, :k].T, b["T"][:, :k] @ b["Q"][:, :k].T) < TOL[precision]
            assert rel(R[:, :k] @ Q[:, :k].T, b["R"][:, :k] @ b["Q"][:, :k].T) < TOL[precision]
    fresh = fp.SimplsBatch(precision, groups, fits, **kw)
    assert fresh.splits() == rep["splits"]
    fresh.close()
    batch.close()
    fp.set_batch_splits(builtin)
    assert fp.batch_splits() == builtin


def test_autotune_rejects_bad_requests(hip):
    fp = hip[0]
    X, Y = latent_data(300, 40, 1, seed=5)
    xb = _dev(fp, X); yb = _dev(fp, Y)
    batch = fp.SimplsBatch(64, [(xb.ptr, 300, 40)], [dict(group=0, q=1, components=2, y=yb.ptr)])
    with pytest.raises(fp.HipError, match="trials"):
        batch.autotune(trials=0)
    with pytest.raises(fp.HipError, match="128..4096"):
        fp.set_batch_splits((64, 1024, 512, 256))
    batch.close()
