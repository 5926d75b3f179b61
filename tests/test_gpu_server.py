"""Request server (GPU, through the C ABI).

* Alone equals in the mix: predictions of resident models and models fitted
  while other requests run are bitwise the same as when each request runs
  alone (batch-invariant query splits, fixed-order reductions, one stream per
  model).
* Server predictions agree with the single-model handle path
  (fastpls_hip_predict_path, which the parity tests pin to the oracle) to
  rounding: 1e-12 relative in FP64, 1e-6 in FP32 (the split of the
  reduction over p differs).
* Admission keeps the bytes in use within the budget; a fit larger than the
  budget is rejected; released models free their bytes; errors are reported
  per request.
"""
import numpy as np
import pytest

from conftest import rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu


def _models(fp):
    XA, YA = latent_data(3000, 400, 3, factors=6, noise=0.2, seed=31)
    XB, YB = latent_data(2000, 300, 4, factors=5, noise=0.2, seed=32)
    XC, LC = class_data(2500, 200, 4, latent=4, seed=33, spread=1.0)
    return [
        dict(x=XA.astype(np.float32), y=YA.astype(np.float32), method=fp.SIMPLS, precision=32,
             components=6, oversample=4, power=1, seed=3),
        dict(x=XB, y=YB, method=fp.PLSSVD, precision=64, components=3, oversample=4, power=2,
             seed=5),
        dict(x=XC.astype(np.float32), labels=LC, classes=4, method=fp.SIMPLS, precision=32,
             components=4, oversample=4, power=1, seed=7),
    ]


def _fit(server, spec):
    kw = {k: v for k, v in spec.items() if k not in ("x", "y", "labels", "classes")}
    return server.submit_fit(spec["x"], y=spec.get("y"), labels=spec.get("labels"),
                             classes=spec.get("classes"), **kw)


def _queries(spec, seed):
    rng = np.random.default_rng(seed)
    X = spec["x"]
    out = []
    for rows in (1, 17, 300, 1000):
        idx = rng.integers(0, X.shape[0], rows)
        out.append(np.asfortranarray(X[idx] + rng.standard_normal((rows, X.shape[1])).astype(X.dtype)))
    return out


def test_server_alone_equals_mix(hip):
    fp = hip[0]
    specs = _models(fp)
    srv = fp.Server(8e9, workers=4, max_batch_rows=4096)
    # alone: one request at a time
    models, alone = [], []
    for spec in specs:
        m = _fit(srv, spec)
        srv.poll(m)
        models.append(m)
    queries = [_queries(spec, 40 + i) for i, spec in enumerate(specs)]
    for i, m in enumerate(models):
        res = []
        for xq in queries[i]:
            rid, out = srv.submit_predict(m, xq)
            srv.poll(rid)
            res.append(out.copy())
        alone.append(res)
    exports_alone = [(srv.export(m, "R", (s["x"].shape[1], s["components"])),
                      srv.export(m, "Q", (s["y"].shape[1] if "y" in s else s["classes"],
                                          s["components"])))
                     for m, s in zip(models[:2], specs[:2])]
    # the mix: refits of the first two models and every query at once
    refits = [_fit(srv, specs[0]), _fit(srv, specs[1])]
    pending = []
    for j in range(4):
        for i, m in enumerate(models):
            pending.append((i, j) + srv.submit_predict(m, queries[i][j]))
    for i, j, rid, out in pending:
        srv.poll(rid)
        np.testing.assert_array_equal(out, alone[i][j])
    for m, (R, Q), s in zip(refits, exports_alone, specs[:2]):
        srv.poll(m)
        np.testing.assert_array_equal(srv.export(m, "R", R.shape), R)
        np.testing.assert_array_equal(srv.export(m, "Q", Q.shape), Q)
    # server predictions agree with the single-model handle path to rounding
    for i, spec in enumerate(specs):
        h = fp.ResidentModel.fit(spec["x"], spec.get("y"), labels=spec.get("labels"),
                                 classes=spec.get("classes"), method=spec["method"],
                                 precision=spec["precision"], components=spec["components"],
                                 oversample=spec["oversample"], power=spec["power"],
                                 seed=spec["seed"])
        for j, xq in enumerate(queries[i]):
            ref = h.predict_path(xq, [spec["components"]])[:, :, 0]
            assert rel(alone[i][j], ref) < (1e-12 if spec["precision"] == 64 else 1e-6), (i, j)
        h.close()
    st = srv.stats()
    assert st["failed"] == 0
    assert st["completed"] == len(specs) + 12 + 12 + 2
    assert st["peak_bytes"] <= st["budget_bytes"]
    assert st["predict_p99_ms"] >= st["predict_p50_ms"] > 0
    srv.close()


def test_server_budget_and_release(hip):
    fp = hip[0]
    X, Y = latent_data(20000, 2000, 2, factors=5, seed=34)
    X32 = X.astype(np.float32)
    Y32 = Y.astype(np.float32)
    # room for one fit of this size at a time (admission estimate ~0.3 GB)
    srv = fp.Server(0.45e9, workers=4, max_batch_rows=2048)
    ids = [srv.submit_fit(X32, Y32, precision=32, components=4, oversample=2, power=1, seed=s)
           for s in (1, 2, 3)]
    for rid in ids:
        srv.poll(rid)
    st = srv.stats()
    assert st["peak_bytes"] <= st["budget_bytes"]
    assert st["peak_concurrent"] == 1
    assert st["resident_models"] == 3
    before = st["bytes_in_use"]
    srv.release(ids[0])
    assert srv.stats()["bytes_in_use"] < before
    with pytest.raises(fp.HipError, match="unknown or released"):
        srv.submit_predict(ids[0], X32[:5])
    with pytest.raises(fp.HipError, match="exceeds the server HBM budget"):
        big = fp.Server(1e8, workers=1)
        big.submit_fit(X32, Y32, precision=32, components=4)
    srv.close()


def test_server_request_errors(hip):
    fp = hip[0]
    X, lab = class_data(300, 20, 3, latent=2, seed=35)
    srv = fp.Server(2e9, workers=2)
    bad = lab.copy()
    bad[3] = 9
    rid = srv.submit_fit(X, labels=bad, classes=3, components=2)
    with pytest.raises(fp.HipError, match="labels"):
        srv.poll(rid)
    m = srv.submit_fit(X, labels=lab, classes=3, components=2)
    srv.poll(m)
    with pytest.raises(fp.HipError, match="prefix exceeds"):
        srv.submit_predict(m, X[:4], prefix=3)
    rid, out = srv.submit_predict(m, X[:4], prefix=2)
    state, latency = srv.poll(rid)
    assert state == fp.Server.DONE and latency > 0 and np.all(np.isfinite(out))
    srv.close()
