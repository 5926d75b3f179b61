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
    args = ap.parse_args()
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
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
