"""A reduced batched-SIMPLS set shaped like the 256-fit benchmark set, checked
fit by fit against the FP64 oracle (GPU, through the C ABI).

The set holds, as the benchmark does, cross-validation folds whose training
rows are shared by several refits with row permutations of the responses
(one unpermuted fit per fold besides), groups with p > n, and uneven
singletons (q from 1 to 8, 2 to 30 components). Every fit is compared
with oracle.simpls_fit on the same training rows and the same permuted
responses (the permutation is applied on the host: Y_used = Y[perm]), with
the batch's per-fit seed.

Precision 32 fits FP32-valued data (the oracle fits the same values in
FP64). Tolerances, at every prefix k of every fit, on the fitted values
X_s B_k = T_k Q_kᵀ and the coefficients B_k = R_k Q_kᵀ (relative Frobenius):
1e-9 (precision 64) and 1e-5 (precision 32), or, where larger, 64 times the
oracle's own FP64 self-consistency at that prefix. That bound is measured
here: the oracle refits with its cross-product S0 perturbed by one unit in
the last place (relative 2^-52 Gaussian noise) and the largest change of B_k
(of X_s B_k) over two draws is taken. Past the latent rank of the data (10
factors) SIMPLS deflates S to a residual of rounding size, and one FP64 ulp
of S0 then moves B_k of these fits by 3e-8 (q = 3, 12 components) up to 1e-4
(q = 1, 30 components): no FP64 implementation with another summation order
can hold B there to 1e-9 of the oracle, so the test holds it to the bound
the oracle itself shows. Within the rank the bound is far below the stated
tolerance and the stated tolerance applies.
"""
import numpy as np
import pytest

from conftest import rel
from oracle.oracle import latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-9, 32: 1e-5}
OVERSAMPLE, POWER = 10, 2
MARGIN = 64.0


def self_consistency(oracle, X, Y, A, seed):
    """Largest change of B_k and of X_s B_k (k = 1..A) when the oracle refits
    with S0 perturbed by relative 2^-52 Gaussian noise (two draws)."""
    pr = oracle.prepare(X, Y, scaling=2)
    xs, S0 = pr["xs"], pr["S"]
    base = oracle.simpls_core(xs, S0, A, oversample=OVERSAMPLE, power=POWER, seed=seed)
    sb, sf = np.zeros(A), np.zeros(A)
    rng = np.random.default_rng(seed)
    for _ in range(2):
        P = oracle.simpls_core(xs, S0 * (1 + 2.0 ** -52 * rng.standard_normal(S0.shape)), A,
                               oversample=OVERSAMPLE, power=POWER, seed=seed)
        for k in range(1, A + 1):
            sb[k - 1] = max(sb[k - 1], rel(P["R"][:, :k] @ P["Q"][:, :k].T,
                                           base["R"][:, :k] @ base["Q"][:, :k].T))
            sf[k - 1] = max(sf[k - 1], rel(P["T"][:, :k] @ P["Q"][:, :k].T,
                                           base["T"][:, :k] @ base["Q"][:, :k].T))
    return sb, sf


def _dev(fp, a):
    return fp.DeviceBuffer.from_array(np.asfortranarray(a))


def c3_set(rng):
    """(X, Y, [(perm or None, A, seed)]) groups of a reduced benchmark-shaped set."""
    groups = []
    seed = 1000
    # two CV data sets: folds x (1 unpermuted + 3 permuted refits)
    for n, p, q, A, nfold in [(900, 240, 3, 12, 3), (420, 700, 4, 9, 2)]:
        X, Y = latent_data(n, p, q, factors=10, noise=0.1, seed=n + p)
        folds = np.arange(n) % nfold
        rng.shuffle(folds)
        for f in range(nfold):
            Xf, Yf = X[folds != f], Y[folds != f]
            fits = []
            for r in range(4):
                perm = None if r == 0 else rng.permutation(Xf.shape[0]).astype(np.int32)
                fits.append((perm, A, seed)); seed += 1
            groups.append((Xf, Yf, fits))
    # uneven singletons, p > n among them
    for n, p, q, A in [(500, 1300, 1, 14), (2400, 180, 8, 25), (350, 90, 1, 30),
                       (1200, 640, 5, 20), (300, 900, 2, 2)]:
        X, Y = latent_data(n, p, q, factors=10, noise=0.1, seed=7 * n + p)
        groups.append((X, Y, [(None, A, seed)])); seed += 1
    return groups


@pytest.mark.parametrize("precision", [64, 32])
def test_c3_shaped_batch_matches_oracle(hip, oracle, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    rng = np.random.default_rng(31)
    keep, groups, fits, ref = [], [], [], []
    for X, Y, members in c3_set(rng):
        Xd, Yd = X.astype(dt), Y.astype(dt)
        xb, yb = _dev(fp, Xd), _dev(fp, Yd)
        keep += [xb, yb]
        groups.append((xb.ptr, X.shape[0], X.shape[1]))
        g = len(groups) - 1
        for perm, A, seed in members:
            f = dict(group=g, q=Y.shape[1], components=A, seed=seed, y=yb.ptr)
            Yu = Yd
            if perm is not None:
                pb = _dev(fp, perm); keep.append(pb)
                f["perm"] = pb.ptr
                Yu = Yd[perm]
            fits.append(f)
            ref.append((Xd.astype(np.float64), Yu.astype(np.float64), A, seed))
    assert any(x.shape[1] > x.shape[0] for x, *_ in ref)
    batch = fp.SimplsBatch(precision, groups, fits, scaling=2, oversample=OVERSAMPLE, power=POWER)
    done, flags, _ = batch.status()
    assert all(fl == 0 for fl in flags), flags
    worst = {"fit": 0.0, "B": 0.0}
    for i, (X, Y, A, seed) in enumerate(ref):
        print(f"fit {i}: n={X.shape[0]} p={X.shape[1]} q={Y.shape[1]} A={A}", flush=True)
        assert done[i] == A, (i, done[i], A)
        o = oracle.simpls_fit(X, Y, ncomp=A, scaling=2, oversample=OVERSAMPLE, power=POWER,
                              seed=seed)
        sb, sf = self_consistency(oracle, X, Y, A, seed)
        # the FP64 values the fit holds: an FP32 export rounds R and Q, which
        # B = R Qᵀ of an ill-conditioned fit magnifies
        Rb, Qb, Tb = (batch.export(i, f, f64=True) for f in ("R", "Q", "T"))
        for k in range(1, A + 1):
            e_fit = rel(Tb[:, :k] @ Qb[:, :k].T, o["T"][:, :k] @ o["Q"][:, :k].T)
            e_b = rel(Rb[:, :k] @ Qb[:, :k].T, o["R"][:, :k] @ o["Q"][:, :k].T)
            tol_fit = max(TOL[precision], MARGIN * sf[k - 1])
            tol_b = max(TOL[precision], MARGIN * sb[k - 1])
            worst["fit"] = max(worst["fit"], e_fit)
            worst["B"] = max(worst["B"], e_b)
            if tol_b == TOL[precision]:
                worst["B_stated"] = max(worst.get("B_stated", 0.0), e_b)
            assert e_fit < tol_fit, (precision, i, k, "fitted", e_fit, tol_fit)
            assert e_b < tol_b, (precision, i, k, "B", e_b, tol_b)
        np.testing.assert_allclose(batch.export(i, "mX", f64=True), o["mx"], rtol=0,
                                   atol=1e-12 * float(np.max(np.abs(o["sx"]))))
        # the batch-precision export is the same values rounded to it
        np.testing.assert_array_equal(batch.export(i, "mX"),
                                      batch.export(i, "mX", f64=True).astype(dt))
    print(f"precision {precision}: worst fitted {worst['fit']:.2e}, worst B {worst['B']:.2e}, "
          f"worst B where the stated tolerance applies {worst.get('B_stated', 0.0):.2e}",
          flush=True)
    batch.close()
