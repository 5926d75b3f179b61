"""Benchmark of the HIP backend on one MI355X.

Workload (default): PLS-SVD through the randomized SVD of XᵀY,
n=100000, p=20000, q=50, 20 components, oversample 10, 2 power iterations,
rSVD seed 1; latent-factor data (30 factors, column norms 0.8^j, noise 0.05,
data seed 0) generated directly in HBM. X is FP32 (8 GB) for the mixed path
(bf16x6 matrix-core products, FP32 accumulation, FP64 small math) and FP64
(16 GB) for the plain path (FP64 throughout). A step is one complete fit
(preprocessing, cross-product, randomized SVD, scores, prefix weights) with
inputs already resident in HBM; value = fits per second of the mixed path.

Accuracy of mixed against plain is reported on identical inputs (the plain
path's FP64 buffers hold the FP32 values): Ŷ on the first 4096 training rows,
and B through 4096 random probes x = mu + sigma z (z ~ N(0, I)):
|Ŷ32(x) - Ŷ64(x)| / |Ŷ64(x) - mu_y| (Frobenius norms).

Prints ONE JSON line on stdout; progress goes to stderr.
`--calibrate-copy MIB` instead times one device-to-device copy of MIB MiB
(the reference kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE) and exits.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Hardware queues per process (HIP's default is 4): the serving section runs
# many streams at once, and streams beyond the queue count share queues.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lib_call(name, *args):
    from fastpls_amd.resident import _err, _raise, load_library
    e = _err()
    if getattr(load_library(), name)(*args, e, 1024) != 0:
        _raise(e)


def gather_rows(buf, n, p, rows, dtype):
    """First `rows` rows of a column-major device matrix, to the host."""
    item = np.dtype(dtype).itemsize
    out = np.empty((rows, p), dtype=dtype, order="F")
    base = out.ctypes.data
    for j in range(p):
        lib_call("fastpls_hip_memcpy", C.c_void_p(base + j * rows * item),
                 C.c_void_p(buf.ptr + j * n * item), C.c_size_t(rows * item), 1)
    return out


def rel(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_fits(fp, xbuf, ybuf, n, p, q, precision, k, steps, warmup, label):
    lat, phases = [], []
    model = None
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        m = fp.ResidentModel.fit_device(xbuf.ptr, ybuf.ptr, None, n, p, q, method=fp.PLSSVD,
                                        precision=precision, components=k, scaling=2,
                                        oversample=10, power=2, seed=1)
        t1 = time.perf_counter()
        ph = m.fit_phases()
        if i >= warmup:
            lat.append(t1 - t0)
            phases.append(ph)
        log(f"[{label}] fit {i + 1}/{warmup + steps}: wall {1e3 * (t1 - t0):.2f} ms, "
            f"device {1e3 * ph['fit']:.2f} ms, cross-product pass {1e3 * ph['crossprod_pass']:.3f} ms, "
            f"score pass {1e3 * ph['score_pass']:.3f} ms")
        if model is not None:
            model.close()
        model = m
    return model, np.array(lat), phases


def cpu_baseline(rows, p, q, k):
    from oracle.oracle import Oracle, latent_data
    orc = Oracle()
    X, Y = latent_data(rows, p, q, factors=30, noise=0.05, seed=0, decay=0.8)
    t0 = time.perf_counter()
    orc.plssvd_fit(X, Y, ncomp=k, oversample=10, power=2, seed=1)
    return time.perf_counter() - t0


def calibrate_copy(fp, mib):
    nbytes = int(mib) << 20
    a, b = fp.DeviceBuffer(nbytes), fp.DeviceBuffer(nbytes)
    lib_call("fastpls_hip_memcpy", C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(nbytes), 2)
    t0 = time.perf_counter()
    lib_call("fastpls_hip_memcpy", C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(nbytes), 2)
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_bytes": nbytes, "seconds": dt, "GBps_read_plus_write": 2 * nbytes / dt / 1e9}),
          flush=True)


def c3_batched(fp, seed=2026, details_out=None):
    """Batched SIMPLS: 256 fits = 2 datasets x 5-fold CV x 20 row permutations
    of the training responses (the fits of one fold share X) + 56 uneven
    singletons; per dataset n ~ U[500, 20000], p ~ U[100, 5000],
    q ~ U[1, 10], ncomp ~ U[2, 30]; latent-factor data (10 factors, noise
    0.1); per-fit rSVD seed 1000 + i. Times the whole batch in FP32 (mixed)
    and in FP64 (plain, the same values) and, as the baseline, the plain FP64
    fits one at a time through the single-fit entry (device-resident
    inputs). Inputs are staged before the timed regions."""
    rng = np.random.default_rng(seed)
    keep, groups32, groups64, fits32, fits64, seq = [], [], [], [], [], []
    for d in range(58):
        n = int(rng.integers(500, 20001)); p = int(rng.integers(100, 5001))
        q = int(rng.integers(1, 11)); A = int(rng.integers(2, 31))
        xb = fp.DeviceBuffer(n * p * 4); yb = fp.DeviceBuffer(n * q * 4)
        fp.latent_data_device(xb.ptr, yb.ptr, 32, n, p, q, factors=10, noise=0.1, seed=100 + d)
        X = xb.download((n, p), np.float32); Y = yb.download((n, q), np.float32)
        xb.free(); yb.free()
        parts = []
        if d < 2:
            folds = np.arange(n) % 5
            rng.shuffle(folds)
            parts = [(X[folds != f], Y[folds != f], 20) for f in range(5)]
        else:
            parts = [(X, Y, 0)]
        for Xg, Yg, nperm in parts:
            ng = Xg.shape[0]
            bufs = [fp.DeviceBuffer.from_array(a) for a in
                    (Xg, Xg.astype(np.float64), Yg, Yg.astype(np.float64))]
            keep += bufs
            groups32.append((bufs[0].ptr, ng, p)); groups64.append((bufs[1].ptr, ng, p))
            g = len(groups32) - 1
            for r in range(max(nperm, 1)):
                i = len(fits32)
                base = dict(group=g, q=q, components=A, seed=1000 + i)
                ys = Yg
                if nperm:
                    perm = rng.permutation(ng).astype(np.int32)
                    pb = fp.DeviceBuffer.from_array(perm); keep.append(pb)
                    base["perm"] = pb.ptr
                    ys = Yg[perm]
                fits32.append(dict(base, y=bufs[2].ptr))
                fits64.append(dict(base, y=bufs[3].ptr))
                yseq = fp.DeviceBuffer.from_array(ys.astype(np.float64)); keep.append(yseq)
                seq.append((bufs[1].ptr, yseq.ptr, ng, p, q, A, 1000 + i, nperm > 0))
    assert len(fits32) == 256
    log(f"[batched] 256 fits over {len(groups32)} shared-X groups staged")

    def run_batch(precision, groups, fits, reorth=-1):
        t0 = time.perf_counter()
        b = fp.SimplsBatch(precision, groups, fits, scaling=2, oversample=10, power=2,
                           reorth=reorth)
        wall = time.perf_counter() - t0
        done, flags, dev = b.status()
        assert all(f == 0 for f in flags) and all(d == f["components"] for d, f in zip(done, fits))
        return b, wall, dev

    # warm-up, then timed batches
    b32, _, _ = run_batch(32, groups32, fits32); b32.close()
    b32, wall32, dev32 = run_batch(32, groups32, fits32)
    log(f"[batched] mixed: {1e3 * wall32:.1f} ms wall, {1e3 * dev32:.1f} ms device")
    b64, wall64, dev64 = run_batch(64, groups64, fits64)
    log(f"[batched] plain: {1e3 * wall64:.1f} ms wall, {1e3 * dev64:.1f} ms device")
    # mixed against plain on the same values, for every fit (the permuted
    # refits included) at every prefix k <= ncomp: fitted values T_k Q_kᵀ and
    # coefficients B_k = R_k Q_kᵀ (relative Frobenius errors)
    err = dict(fitted=0.0, b=0.0, fitted_rank=0.0, b_rank=0.0)
    details = []
    for i, f in enumerate(fits32):
        T32, Q32 = b32.export(i, "T").astype(np.float64), b32.export(i, "Q").astype(np.float64)
        R32 = b32.export(i, "R").astype(np.float64)
        T64, Q64, R64 = b64.export(i, "T"), b64.export(i, "Q"), b64.export(i, "R")
        A = f["components"]
        e_fit = [rel(T32[:, :k] @ Q32[:, :k].T, T64[:, :k] @ Q64[:, :k].T) for k in range(1, A + 1)]
        e_b = [rel(R32[:, :k] @ Q32[:, :k].T, R64[:, :k] @ Q64[:, :k].T) for k in range(1, A + 1)]
        kr = min(A, 10)
        e = dict(fitted=max(e_fit), b=max(e_b), fitted_rank=e_fit[kr - 1], b_rank=e_b[kr - 1])
        for key, v in e.items():
            err[key] = max(err[key], v)
        g = groups32[f["group"]]
        details.append(dict(fit=i, n=g[1], p=g[2], q=f["q"], A=A, permuted="perm" in f, **e))
    b32.close(); b64.close()
    if details_out is not None:
        details_out.extend(details)
    # baseline: plain FP64, one fit at a time
    lat = []
    for (xp, yp, n, p, q, A, sd, _) in seq:
        t0 = time.perf_counter()
        m = fp.ResidentModel.fit_device(xp, yp, None, n, p, q, method=fp.SIMPLS, precision=64,
                                        components=A, scaling=2, oversample=10, power=2, seed=sd)
        lat.append(time.perf_counter() - t0)
        m.close()
    lat = np.array(lat)
    log(f"[batched] one at a time (plain FP64): {1e3 * lat.sum():.1f} ms")
    for b in keep:
        b.free()
    return {
        "workload": "256 SIMPLS fits: 2 x (5-fold CV x 20 response-row permutations) + 56 "
                    "singletons; n~U[500,2e4], p~U[100,5e3], q~U[1,10], ncomp~U[2,30]; "
                    "oversample 10, power 2",
        "batched_mixed_fits_per_s": round(256 / wall32, 2),
        "batched_mixed_wall_ms": round(1e3 * wall32, 2),
        "batched_plain_fits_per_s": round(256 / wall64, 2),
        "sequential_plain_fits_per_s": round(256 / float(lat.sum()), 2),
        "sequential_plain_p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "sequential_plain_p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "speedup_batched_mixed_vs_sequential_plain": round(float(lat.sum()) / wall32, 2),
        "speedup_batched_mixed_vs_batched_plain": round(wall64 / wall32, 2),
        "max_rel_err_fitted_mixed_vs_plain_at_rank": err["fitted_rank"],
        "max_rel_err_B_mixed_vs_plain_at_rank": err["b_rank"],
        "max_rel_err_fitted_mixed_vs_plain_all_components": err["fitted"],
        "max_rel_err_B_mixed_vs_plain_all_components": err["b"],
        "accuracy_gate_1e-5_met": bool(max(err.values()) <= 1e-5),
        "accuracy_note": "maxima over all 256 fits (permuted refits included) and every prefix "
                         "k <= ncomp ('at rank': k = min(ncomp, 10), the data's latent rank); "
                         "relative Frobenius errors of T_k Q_k' and R_k Q_k' against the plain "
                         "FP64 batch on identical values",
    }


def c5_serving(fp, requests=10000, rate=500.0, seed=3, workers=32, max_fits=0):
    """Serving stream: `requests` requests with Poisson arrivals at `rate`/s
    (seed 3) against the request server (HBM budget 200 GB, FP32 storage,
    bf16x6 matrix-core products with FP32 accumulation). 80% predict 1..4096
    rows against one of 16 resident SIMPLS models (p ~ U[1000, 20000],
    q ~ U[1, 20], 10 components, fitted from 4096 latent-factor rows); 20%
    fit a SIMPLS model (n log-U[1e3, 1e5], p log-U[1e2, 2e4], ncomp
    U{2..25}, q U{1..20}) on one of 32 latent-factor data sets drawn from
    that size distribution, and the fitted model is released once done.
    Predict rows and fit data are resident in HBM before the stream starts
    (device inputs); latency runs from submission to completion."""
    rng = np.random.default_rng(seed)
    keep = []
    srv = fp.Server(200e9, workers=workers, max_batch_rows=8192, max_concurrent_fits=max_fits)
    models = []
    for i in range(16):
        p = int(rng.integers(1000, 20001)); q = int(rng.integers(1, 21)); n = 4096
        xb = fp.DeviceBuffer(n * p * 4); yb = fp.DeviceBuffer(n * q * 4)
        fp.latent_data_device(xb.ptr, yb.ptr, 32, n, p, q, factors=10, noise=0.1, seed=500 + i)
        mid = srv.submit_fit(None, precision=32, components=10, oversample=10, power=2, seed=1,
                             device=(xb.ptr, yb.ptr, n, p, q))
        srv.poll(mid)
        qb = fp.DeviceBuffer(4096 * p * 4); qy = fp.DeviceBuffer(4096 * q * 4)
        fp.latent_data_device(qb.ptr, qy.ptr, 32, 4096, p, q, factors=10, noise=0.1, seed=900 + i)
        keep += [xb, yb, qb, qy]
        models.append((mid, qb))
    pool = []
    for d in range(32):
        n = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e5)))))
        p = int(round(np.exp(rng.uniform(np.log(1e2), np.log(2e4)))))
        xb = fp.DeviceBuffer(n * p * 4); yb = fp.DeviceBuffer(n * 20 * 4)
        fp.latent_data_device(xb.ptr, yb.ptr, 32, n, p, 20, factors=10, noise=0.1, seed=700 + d)
        keep += [xb, yb]
        pool.append((xb, yb, n, p))
    data_gb = sum(b.nbytes for b in keep) / 1e9
    log(f"[serving] 16 resident models and 32 fit data sets staged ({data_gb:.1f} GB)")
    arrivals = np.cumsum(rng.exponential(1.0 / rate, requests))
    is_fit = rng.random(requests) < 0.2
    subs = []
    t0 = time.perf_counter()
    for i in range(requests):
        while True:
            wait = arrivals[i] - (time.perf_counter() - t0)
            if wait <= 0:
                break
            time.sleep(min(wait, 0.002))
        if is_fit[i]:
            xb, yb, n, p = pool[int(rng.integers(0, 32))]
            A = min(int(rng.integers(2, 26)), p, n - 1)
            q = int(rng.integers(1, 21))
            rid = srv.submit_fit(None, precision=32, components=A, oversample=10, power=2,
                                 seed=1 + i, device=(xb.ptr, yb.ptr, n, p, q))
            subs.append((True, rid))
        else:
            mid, qb = models[int(rng.integers(0, 16))]
            rows = int(rng.integers(1, 4097))
            rid, _ = srv.submit_predict(mid, device=(qb.ptr, 4096, rows))
            subs.append((False, rid))
        if (i + 1) % 2000 == 0:
            log(f"[serving] {i + 1}/{requests} submitted")
    for fit, rid in subs:
        srv.poll(rid)
        if fit:
            srv.release(rid)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    for b in keep:
        b.free()
    log(f"[serving] {requests} requests in {wall:.2f} s: predict p50 {st['predict_p50_ms']:.2f} ms "
        f"p99 {st['predict_p99_ms']:.2f} ms, fit p50 {st['fit_p50_ms']:.2f} ms "
        f"p99 {st['fit_p99_ms']:.2f} ms")
    return {
        "workload": f"{requests} requests, Poisson {rate:g}/s (seed {seed}); 80% predict "
                    "(1..4096 rows, 16 resident SIMPLS models, p~U[1e3,2e4], q~U[1,20]); 20% "
                    "SIMPLS fit (n log-U[1e3,1e5], p log-U[1e2,2e4], ncomp U{2..25}, q U{1..20}, "
                    "32 staged data sets); HBM budget 200 GB; FP32 storage, bf16x6 MFMA",
        "offered_per_s": rate,
        "completed_per_s": round(requests / wall, 2),
        "predict_p50_ms": round(st["predict_p50_ms"], 3),
        "predict_p99_ms": round(st["predict_p99_ms"], 3),
        "fit_p50_ms": round(st["fit_p50_ms"], 3),
        "fit_p99_ms": round(st["fit_p99_ms"], 3),
        "failed": int(st["failed"]),
        "peak_hbm_gb_admitted": round(st["peak_bytes"] / 1e9, 3),
        "budget_gb": 200,
        "peak_concurrent_batches": int(st["peak_concurrent"]),
        "predict_batches": int(st["predict_batches"]),
        "mean_rows_per_predict_batch": round(st["mean_batch_rows"], 1),
        "max_requests_per_batch": int(st["max_batch_requests"]),
        "fit_service_ms_mean": round(st["fit_service_ms"], 3),
        "predict_service_ms_mean": round(st["predict_service_ms"], 3),
        "workers": workers,
        "max_concurrent_fits": max_fits if max_fits > 0 else max(1, workers // 2),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--p", type=int, default=20000)
    ap.add_argument("--q", type=int, default=50)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--cpu-rows", type=int, default=6000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=10000)
    args = ap.parse_args()
    if args.steps < 1 or args.plain_steps < 1:
        ap.error("--steps and --plain-steps need at least one fit (use --no-plain to skip the FP64 line)")
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent fits would shard one scheduler per GPU).", file=sys.stderr)
        sys.exit(2)

    import fastpls_amd as fp
    if args.calibrate_copy:
        calibrate_copy(fp, args.calibrate_copy)
        return
    info = fp.device_info()
    log(f"device: {info}")
    n, p, q, k = args.n, args.p, args.q, args.k
    rows = min(4096, n)

    # ---------------------------------------------------- mixed (FP32 X)
    log("generating FP32 latent-factor data in HBM")
    x32 = fp.DeviceBuffer(n * p * 4)
    y32 = fp.DeviceBuffer(n * q * 4)
    fp.latent_data_device(x32.ptr, y32.ptr, 32, n, p, q, factors=30, decay=0.8, noise=0.05, seed=0)
    m32, lat32, ph32 = run_fits(fp, x32, y32, n, p, q, 32, k, args.steps, args.warmup, "mixed")
    wall32 = float(lat32.sum())
    fps = args.steps / wall32
    xr32 = gather_rows(x32, n, p, rows, np.float32)
    x32.free(); y32.free()

    # ---------------------------------------------------- plain (FP64 X)
    plain = {}
    if not args.no_plain:
        # the same data as the mixed path (FP32 values held in FP64), so the
        # comparison measures the mixed path's arithmetic alone
        log("generating the same latent-factor data in FP64 storage")
        x64 = fp.DeviceBuffer(n * p * 8)
        y64 = fp.DeviceBuffer(n * q * 8)
        fp.latent_data_device(x64.ptr, y64.ptr, 64, n, p, q, factors=30, decay=0.8, noise=0.05,
                              seed=0, round32=True)
        m64, lat64, ph64 = run_fits(fp, x64, y64, n, p, q, 64, k, args.plain_steps, 1, "plain")
        x64.free(); y64.free()
        yhat32 = m32.predict_path(xr32, [k]).astype(np.float64)[:, :, 0]
        yhat64 = m64.predict_path(xr32.astype(np.float64), [k])[:, :, 0]
        mu, sd, my = m64.export("mX"), m64.export("vX"), m64.export("mY")
        z = np.random.default_rng(7).standard_normal((rows, p))
        probe = np.asfortranarray(mu[None, :] + sd[None, :] * z)
        b32 = m32.predict_path(probe.astype(np.float32), [k]).astype(np.float64)[:, :, 0]
        b64 = m64.predict_path(probe, [k])[:, :, 0]
        plain = {
            "plain_fits_per_s": round(args.plain_steps / float(lat64.sum()), 4),
            "plain_ms_per_fit": round(1e3 * float(np.median(lat64)), 3),
            "speedup_mixed_vs_plain": round(float(np.median(lat64)) / float(np.median(lat32)), 3),
            "rel_err_yhat_mixed_vs_plain": rel(yhat32, yhat64),
            "rel_err_B_probe_mixed_vs_plain": float(np.linalg.norm(b32 - b64) /
                                                    np.linalg.norm(b64 - my[None, :])),
        }
        m64.close()
    m32.close()

    # ------------------------------------------------- roofline (X passes)
    # Each event interval spans the operand split kernel (tens of µs) and the
    # matrix-core pass that reads X once: algorithmic bytes = n p 4 (X) +
    # n q 4 (Y) for the cross-product pass and n p 4 + n k 8 for the scores.
    cp = np.array([ph["crossprod_pass"] for ph in ph32])
    sc = np.array([ph["score_pass"] for ph in ph32])
    cp_bytes = float(n) * p * 4 + float(n) * q * 4
    sc_bytes = float(n) * p * 4 + float(n) * k * 8
    t_cp, t_sc = float(np.mean(cp)), float(np.mean(sc))
    achieved = cp_bytes / t_cp
    # HBM bytes per launch of the same kernel from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs (scaled by a copy of known size), when
    # they were taken at this workload
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "plssvd_c2_traffic.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath))
        if (tr.get("kernel") == "k_xtw_bf6<64>" and
                tr.get("workload", "").startswith(f"plssvd_rsvd n={n} p={p} q={q} k={k}")):
            traffic = tr["traffic_bytes"]
    roofline = {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                "algorithmic_bytes": cp_bytes,
                "kernel": "k_xtw_bf6<64> (column stats + X^T Y, one pass over X, bf16x6 MFMA)",
                "kernel_ms": round(1e3 * t_cp, 4),
                "score_pass_kernel": "k_xw_bf6<32> (T = X_s R, bf16x6 MFMA)",
                "score_pass_ms": round(1e3 * t_sc, 4),
                "score_pass_frac": round(sc_bytes / t_sc / HBM_PEAK, 4)}

    # ----------------------------------------------- batched SIMPLS fits
    batched = None
    if not args.no_batched:
        batched = c3_batched(fp)

    # ------------------------------------------------------------ serving
    serving = None
    if not args.no_serving:
        serving = c5_serving(fp, requests=args.serve_requests)

    # ------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        log(f"CPU oracle on a {args.cpu_rows}-row sample (1 thread)")
        dt = cpu_baseline(args.cpu_rows, p, q, k)
        t_full = dt * n / args.cpu_rows
        cpu = {"value": round(1.0 / t_full, 6), "unit": "fits/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle PLS-SVD fit on {args.cpu_rows} of {n} rows (same p, q, k, "
                         f"oversample, power), {dt:.2f} s, scaled linearly in n"}

    out = {
        "metric": "SIMPLS fits/s + p50/p99 latency, mixed vs plain FP64 GPU at rel err <=1e-5",
        "value": round(fps, 4),
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall32 / args.steps, 3),
        "p50_ms": round(1e3 * float(np.percentile(lat32, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat32, 99)), 3),
        "device_ms_per_fit": round(1e3 * float(np.mean([ph["fit"] for ph in ph32])), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage; bf16x6 MFMA with f32 accumulation; f64 small math",
        "data": "synthetic latent-factor data generated in HBM (Philox4x32-10)",
        "config": {"workload": f"plssvd_rsvd n={n} p={p} q={q} k={k} oversample=10 power=2 "
                               f"(X fp32 {float(n) * p * 4 / 1e9:.1f} GB)",
                   "n": n, "p": p, "q": q, "components": k},
        "device": info,
        "roofline": roofline,
        "cpu_baseline": cpu,
        **plain,
        "batched_simpls": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
