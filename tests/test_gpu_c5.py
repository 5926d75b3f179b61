"""A reduced serving stream shaped like the 10 000-request benchmark stream
(GPU, through the C ABI).

300 requests with Poisson arrivals at 500/s (seed 3), paced in real time so
that requests arrive while others run on the device: 80% predict requests of
1..512 rows against one of six resident FP32 SIMPLS models, 20% FP32 SIMPLS
fit requests on one of eight data sets (n log-uniform in [500, 5000], p
log-uniform in [50, 1500], q in 1..4, 2..8 components, within the data's
latent rank of 8). Checked:

* every predict result is bitwise the result of the same request replayed
  alone on the idle server afterwards (continuous batching does not change
  results);
* every predict agrees with the single-model handle path
  (fastpls_hip_predict_path of a model fitted alone with the same inputs)
  to 1e-6 relative (FP32 outputs; the split of the reduction over p
  differs);
* every model fitted in the stream matches the FP64 oracle on the same FP32
  values: B = R Q' at every prefix to 1e-5 relative (R and Q are exported in
  FP32), and a few of them are bitwise their refit alone;
* nothing failed and every request completed.
"""
import time

import numpy as np
import pytest

from conftest import rel
from oracle.oracle import latent_data

pytestmark = pytest.mark.gpu


def test_c5_shaped_stream(hip, oracle):
    fp = hip[0]
    rng = np.random.default_rng(3)
    srv = fp.Server(20e9, workers=8, max_batch_rows=2048)
    # six resident models
    models = []
    for i in range(6):
        p = int(rng.integers(300, 2001)); q = int(rng.integers(1, 7))
        X, Y = latent_data(1500, p, q, factors=8, noise=0.1, seed=500 + i)
        X, Y = X.astype(np.float32), Y.astype(np.float32)
        mid = srv.submit_fit(X, Y, precision=32, components=6, oversample=4, power=2, seed=1)
        srv.poll(mid)
        Xq, _ = latent_data(512, p, q, factors=8, noise=0.1, seed=900 + i)
        models.append(dict(id=mid, X=X, Y=Y, Xq=Xq.astype(np.float32), q=q))
    # eight fit data sets
    pool = []
    for d in range(8):
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(5000)))))
        p = int(round(np.exp(rng.uniform(np.log(50), np.log(1500)))))
        q = int(rng.integers(1, 5))
        X, Y = latent_data(n, p, q, factors=8, noise=0.1, seed=700 + d)
        pool.append((X.astype(np.float32), Y.astype(np.float32)))
    # the stream
    requests = 300
    arrivals = np.cumsum(rng.exponential(1.0 / 500.0, requests))
    is_fit = rng.random(requests) < 0.2
    subs = []
    t0 = time.perf_counter()
    for i in range(requests):
        wait = arrivals[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        if is_fit[i]:
            X, Y = pool[int(rng.integers(0, 8))]
            A = min(int(rng.integers(2, 9)), X.shape[1], X.shape[0] - 1)
            rid = srv.submit_fit(X, Y, precision=32, components=A, oversample=4, power=2,
                                 seed=1 + i)
            subs.append(("fit", rid, X, Y, A, 1 + i))
        else:
            m = models[int(rng.integers(0, 6))]
            rows = int(rng.integers(1, 513))
            prefix = int(rng.integers(1, 7))
            xq = np.asfortranarray(m["Xq"][:rows])
            rid, out = srv.submit_predict(m["id"], xq, prefix=prefix)
            subs.append(("predict", rid, m, xq, prefix, out))
    for s in subs:
        srv.poll(s[1])
    st = srv.stats()
    assert st["failed"] == 0
    assert st["completed"] == requests + 6
    assert st["peak_concurrent"] > 1  # requests did overlap
    # alone replay, single-model handle path
    handles = {}
    for m in models:
        handles[m["id"]] = fp.ResidentModel.fit(m["X"], m["Y"], precision=32, components=6,
                                                oversample=4, power=2, seed=1)
    npred = 0
    for kind, rid, *rest in subs:
        if kind != "predict":
            continue
        m, xq, prefix, out = rest
        rid2, alone = srv.submit_predict(m["id"], xq, prefix=prefix)
        srv.poll(rid2)
        np.testing.assert_array_equal(out, alone)
        ref = handles[m["id"]].predict(xq, prefix)
        assert rel(out, ref) < 1e-6, (rid, rel(out, ref))
        npred += 1
    for h in handles.values():
        h.close()
    assert npred > 200
    # fitted models against the oracle; a few refitted alone, bitwise
    refits = 0
    for kind, rid, *rest in subs:
        if kind != "fit":
            continue
        X, Y, A, seed = rest
        R = srv.export(rid, "R", (X.shape[1], A)).astype(np.float64)
        Q = srv.export(rid, "Q", (Y.shape[1], A)).astype(np.float64)
        o = oracle.simpls_fit(X.astype(np.float64), Y.astype(np.float64), ncomp=A, scaling=2,
                              oversample=4, power=2, seed=seed)
        for k in range(1, A + 1):
            e = rel(R[:, :k] @ Q[:, :k].T, o["R"][:, :k] @ o["Q"][:, :k].T)
            assert e < 1e-5, (rid, k, e)
        if refits < 3:
            again = srv.submit_fit(X, Y, precision=32, components=A, oversample=4, power=2,
                                   seed=seed)
            srv.poll(again)
            np.testing.assert_array_equal(srv.export(again, "R", (X.shape[1], A)),
                                          R.astype(np.float32))
            srv.release(again)
            refits += 1
        srv.release(rid)
    assert refits == 3
    srv.close()
