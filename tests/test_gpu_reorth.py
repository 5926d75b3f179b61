"""Re-orthogonalising SIMPLS (the reference float core's routing) on the GPU.

The reference re-orthogonalises when its arithmetic is FP32
(src/r_api.cpp:1416) and on the rank-one route (src/r_api.cpp:1491): twice
per component the projection of the score on the earlier scores is removed
from it, and the same combination of the earlier weights from the weight
(simpls.hpp:161-200, 635-642); the loading is the product with the corrected
score, and its projection on the earlier deflation basis is removed twice.
The batch entry forces this routing with reorth = 1. In exact arithmetic the
SIMPLS scores are orthogonal and both routings give one model; in floating
point they part past the latent rank: on the data below the oracle's two
routings differ by 2.4e-5 (regression) and 8e-8 (labels) relative in B at
30 components, and the single-pass scores lose orthogonality at the 1e-9
to 1e-7 level. So each fit is
checked against the oracle's re-orthogonalising routing at every prefix,
against the single-pass routing to show that the check separates the two,
and for orthonormal scores.

Tolerances (relative Frobenius): precision 64, 1e-9 on B and the fitted
values T Qᵀ at every prefix; precision 32 (X and Y stored FP32, FP64
products and small math), the same against the oracle on the FP32 values
widened to FP64.
"""
import numpy as np
import pytest

from conftest import rel
from oracle.oracle import class_data, latent_data

pytestmark = pytest.mark.gpu

A = 30


def coef(R, Q, k):
    return R[:, :k].astype(np.float64) @ Q[:, :k].astype(np.float64).T


def _dev(fp, a):
    return fp.DeviceBuffer.from_array(np.asfortranarray(a))


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_reorth_matches_float_core_routing(hip, oracle, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    X, Y = latent_data(1200, 3000, 3, factors=12, noise=0.05, seed=7, decay=0.85)
    XC, lab = class_data(900, 2000, 5, latent=8, seed=5, spread=1.0)
    Xp, Yp, XCp = X.astype(dt), Y.astype(dt), XC.astype(dt)
    keep = [_dev(fp, Xp), _dev(fp, Yp), _dev(fp, XCp), _dev(fp, lab.astype(np.int32))]
    groups = [(keep[0].ptr, 1200, 3000), (keep[2].ptr, 900, 2000)]
    fits = [dict(group=0, q=3, components=A, seed=3, y=keep[1].ptr),
            dict(group=1, q=5, components=A, seed=4, labels=keep[3].ptr)]
    batch = fp.SimplsBatch(precision, groups, fits, scaling=2, oversample=10, power=2, reorth=1)
    done, flags, _ = batch.status()
    assert all(f == 0 for f in flags), flags
    assert list(done) == [A, A], done

    kw = dict(ncomp=A, scaling=2, oversample=10, power=2)
    refs = [
        [oracle.simpls_fit(Xp.astype(np.float64), Yp.astype(np.float64), seed=3, reorth=r, **kw)
         for r in (1, 0)],
        [oracle.simpls_fit(XCp.astype(np.float64), labels=lab, classes=5, seed=4, reorth=r, **kw)
         for r in (1, 0)],
    ]
    # the routings' distance at 30 components, at least (oracle, FP64)
    separation = (1e-6, 1e-8)
    for i, (two, one) in enumerate(refs):
        assert two["completed"] == A and one["completed"] == A
        R, Q, T = (batch.export(i, f, f64=True) for f in ("R", "Q", "T"))
        worst = 0.0
        for k in range(1, A + 1):
            eb = rel(coef(R, Q, k), coef(two["R"], two["Q"], k))
            ef = rel(coef(T, Q, k), coef(two["T"], two["Q"], k))
            worst = max(worst, eb, ef)
            assert eb < 1e-9 and ef < 1e-9, (precision, i, k, eb, ef)
        apart = rel(coef(two["R"], two["Q"], A), coef(one["R"], one["Q"], A))
        got_apart = rel(coef(R, Q, A), coef(one["R"], one["Q"], A))
        print(f"precision {precision} fit {i}: worst {worst:.2e}; routings apart {apart:.2e}, "
              f"GPU from single-pass {got_apart:.2e}")
        assert apart > separation[i] and got_apart > separation[i], (i, apart, got_apart)
        T64 = T.astype(np.float64)
        G = T64.T @ T64
        assert np.abs(G - np.eye(A)).max() < 1e-12, (i, np.abs(G - np.eye(A)).max())
    batch.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_reorth_matches_reference_fixture(hip, precision):
    """The same routing against the compiled reference core's own output
    (`simpls_dense_reorth`: 30 components of a rank-8 problem, 200 x 300 x 3)
    at every prefix; precision 32 stores the inputs in FP32, so it is checked
    on B within 1e-5 over the latent rank and on the fitted values throughout."""
    import hashlib

    from conftest import load_golden
    fp = hip[0]
    g = load_golden("simpls_dense_reorth")
    n, p, q, factors, noise, seed, decay = g["recipe"]
    X, Y = latent_data(int(n), int(p), int(q), factors=int(factors), noise=float(noise),
                       seed=int(seed), decay=float(decay))
    assert hashlib.sha256(np.asfortranarray(X).tobytes(order="F")).hexdigest() == str(g["x_sha256"])
    dt = np.float64 if precision == 64 else np.float32
    A = int(g["ncomp"])
    keep = [_dev(fp, X.astype(dt)), _dev(fp, Y.astype(dt))]
    batch = fp.SimplsBatch(precision, [(keep[0].ptr, int(n), int(p))],
                           [dict(group=0, q=int(q), components=A, seed=int(g["seed"]), y=keep[1].ptr)],
                           scaling=2, oversample=int(g["oversample"]), power=int(g["power"]), reorth=1)
    done, flags, _ = batch.status()
    assert flags[0] == 0 and done[0] == A
    R, Q, T = (batch.export(0, f, f64=True) for f in ("R", "Q", "T"))
    for k in range(1, A + 1):
        ef = rel(coef(T, Q, k), coef(g["T"], g["Q"], k))
        eb = rel(coef(R, Q, k), coef(g["R"], g["Q"], k))
        if precision == 64:
            assert eb < 1e-9 and ef < 1e-9, (k, eb, ef)
        else:
            assert ef < 1e-5, (k, ef)
            if k <= int(factors):
                assert eb < 1e-5, (k, eb)
    batch.close()
