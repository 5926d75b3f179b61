"""Write the golden vectors under tests/golden/ from the reference core.

The reference's own header-only core (inst/include/fastpls/core) and its test
backend (core/tests/reference_backend.hpp) are compiled by ``make -C oracle
ref`` into oracle/_ref/libfastpls_ref.so; this script runs it on seeded
synthetic inputs and stores inputs (or their generator recipe and checksum)
with the outputs as .npz fixtures. Run it where the reference checkout is
present:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import AUDIT_FIELDS, Reference, build, class_data, latent_data  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def simpls_case(ref, name, X, Y=None, labels=None, classes=None, ncomp=5, oversample=32,
                power=5, seed=1, block=1, cache=0, batch_geometry=0, store_inputs=True,
                recipe=None, scaling=2):
    m = ref.simpls_fit(X, Y, labels=labels, classes=classes, ncomp=ncomp, oversample=oversample,
                       power=power, seed=seed, block=block, cache=cache,
                       batch_geometry=batch_geometry, scaling=scaling)
    arrays = dict(R=m["R"], Q=m["Q"], V=m["V"], T=m["T"], mx=m["mx"], sx=m["sx"], my=m["my"],
                  completed=np.int64(m["completed"]), ncomp=np.int64(ncomp),
                  oversample=np.int64(oversample), power=np.int64(power), seed=np.int64(seed),
                  block=np.int64(block), cache=np.int64(cache),
                  batch_geometry=np.int64(batch_geometry), x_sha256=digest(X))
    if scaling != 2:
        arrays["scaling"] = np.int64(scaling)
    if store_inputs:
        arrays["X"] = X
        if Y is not None:
            arrays["Y"] = Y
    if labels is not None:
        arrays["labels"] = np.asarray(labels, dtype=np.int32)
        arrays["classes"] = np.int64(classes)
    if recipe is not None:
        arrays["recipe"] = np.asarray(recipe, dtype=np.float64)
    save(name, **arrays)


def plssvd_case(ref, name, X, Y=None, labels=None, classes=None, ncomp=3, oversample=32,
                power=5, seed=1, scaling=2):
    m = ref.plssvd_fit(X, Y, labels=labels, classes=classes, ncomp=ncomp, oversample=oversample,
                       power=power, seed=seed, scaling=scaling)
    arrays = dict(R=m["R"], Q=m["Q"], T=m["T"], sigma=m["sigma"], mx=m["mx"], sx=m["sx"],
                  my=m["my"], ncomp=np.int64(ncomp), oversample=np.int64(oversample),
                  power=np.int64(power), seed=np.int64(seed), X=X)
    if scaling != 2:
        arrays["scaling"] = np.int64(scaling)
    for k, w in enumerate(m["W"]):
        arrays[f"W{k + 1}"] = w
    if Y is not None:
        arrays["Y"] = Y
    if labels is not None:
        arrays["labels"] = np.asarray(labels, dtype=np.int32)
        arrays["classes"] = np.int64(classes)
    save(name, **arrays)


def audited_cases(ref):
    """fastsvd (audited operator rSVD): certified on the first attempt, after
    three attempts, the exact route below six columns, a wide matrix and a
    left-only call."""
    rng = np.random.default_rng(51)
    cases = [
        ("audited_lowrank", latent_data(300, 120, 1, factors=8, noise=0.05, seed=3)[0], 5, 10, 2,
         7, False),
        ("audited_flat", rng.standard_normal((200, 90)), 5, 10, 2, 7, False),
        ("audited_tiny", rng.standard_normal((5, 40)), 3, 32, 5, 7, False),
        ("audited_wide", latent_data(60, 400, 1, factors=6, noise=0.1, seed=4)[0], 8, 0, 0, 9,
         False),
        ("audited_left", latent_data(150, 80, 1, factors=5, noise=0.1, seed=5)[0], 4, 6, 1, 3,
         True),
    ]
    for name, A, k, os_, pw, seed, left in cases:
        r = ref.audited_rsvd(A, k, oversample=os_, power=pw, seed=seed, left_only=left)
        arrays = dict(A=A, U=r["U"], s=r["s"], k=np.int64(k), oversample=np.int64(os_),
                      power=np.int64(pw), seed=np.int64(seed), left_only=np.int64(left),
                      audit=np.array([r["audit"][f] for f in AUDIT_FIELDS]))
        if not left:
            arrays["Vt"] = r["Vt"]
        save(name, **arrays)


def operator_cases(ref):
    """Matrix-free SIMPLS and PLS-SVD (simpls.hpp:768-1047, plssvd.hpp:222-248)
    with the reference's centred cross-covariance and projected operators:
    the block route (block 3, batched geometry), the rank-one direction route
    of a cross-covariance above 512 MiB (block 1, re-orthogonalisation), and
    PLS-SVD."""
    X, Y = latent_data(200, 60, 12, factors=8, noise=0.1, seed=3)
    for name, kw in (("simpls_operator_block", dict(block=3, batch_geometry=1, rank_one=0)),
                     ("simpls_operator_rank1", dict(block=1, reorth=1, rank_one=1))):
        m = ref.simpls_operator_fit(X, Y, ncomp=6, oversample=4, power=3, seed=5, **kw)
        save(name, X=X, Y=Y, R=m["R"], Q=m["Q"], V=m["V"], T=m["T"], mx=m["mx"], sx=m["sx"],
             my=m["my"], completed=np.int64(m["completed"]), ncomp=np.int64(6),
             oversample=np.int64(4), power=np.int64(3), seed=np.int64(5),
             **{k: np.int64(v) for k, v in kw.items()})
    X, Y = latent_data(150, 40, 25, factors=6, noise=0.1, seed=4)
    m = ref.plssvd_operator_fit(X, Y, ncomp=4, oversample=3, power=2, seed=7)
    arrays = dict(X=X, Y=Y, R=m["R"], Q=m["Q"], T=m["T"], sigma=m["sigma"], mx=m["mx"],
                  sx=m["sx"], my=m["my"], ncomp=np.int64(4), oversample=np.int64(3),
                  power=np.int64(2), seed=np.int64(7))
    for k, w in enumerate(m["W"]):
        arrays[f"W{k + 1}"] = w
    save("plssvd_operator", **arrays)


def edge_cases(ref):
    """Scaling 1 (centring, R's default, R/main.R:9746-9748) and 3 (none) for
    SIMPLS (dense and labels) and PLS-SVD; the LDA ridge ladder past 1e-8 and
    its exhaustion (lda.hpp:86-131), and the PLS-SVD prefix solve's pivoted
    fallback and its failure (plssvd.hpp:138-164, linalg.hpp:57-177), from
    exactly constructed moments."""
    X, Y = latent_data(150, 40, 3, factors=5, noise=0.1, seed=61)
    for sc in (1, 3):
        simpls_case(ref, f"simpls_scaling{sc}", X, Y, ncomp=5, seed=7, scaling=sc)
    X, lab = class_data(300, 25, 4, latent=4, seed=62)
    simpls_case(ref, "simpls_labels_scaling1", X, labels=lab, classes=4, ncomp=3, oversample=1,
                power=2, seed=5, scaling=1)
    X, Y = latent_data(120, 30, 8, factors=6, noise=0.1, seed=63)
    for sc in (1, 3):
        plssvd_case(ref, f"plssvd_scaling{sc}", X, Y, ncomp=3, oversample=2, power=2, seed=9,
                    scaling=sc)

    # LDA from moments: a pooled covariance with eigenvalues (2, 1, -6e-6)
    # (rotated) needs the 1e-5 rung at prefix 3; one with eigenvalue -0.5
    # fails at every rung
    rng = np.random.default_rng(64)
    Qr, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    counts = np.array([4, 5, 3], dtype=np.int32)
    n = int(counts.sum())
    means = rng.standard_normal((3, 3))
    sums = means * counts[:, None]

    def gram_of(eig):
        pooled = Qr @ np.diag(eig) @ Qr.T
        return pooled * (n - 3) + sum(counts[c] * np.outer(means[c], means[c]) for c in range(3))
    arrays = dict(sums=sums, counts=counts, n=np.int64(n))
    arrays["gram"] = gram_of([2.0, 1.0, -6e-6])
    arrays["gram_fail"] = gram_of([2.0, 1.0, -0.5])
    for k in (1, 2, 3):
        m = ref.lda_moments(arrays["gram"], sums, counts, n, k)
        arrays[f"linear{k}"] = m["linear"]
        arrays[f"constants{k}"] = m["constants"]
        arrays[f"ridge{k}"] = np.float64(m["ridge"])
        arrays[f"relative{k}"] = np.float64(m["relative"])
    try:
        ref.lda_moments(arrays["gram_fail"], sums, counts, n, 3)
        raise SystemExit("the failing LDA case was solved")
    except RuntimeError:
        pass
    save("lda_ladder", **arrays)

    # PLS-SVD prefix weights: prefix 1 takes Cholesky, prefixes 2-4 the
    # pivoted solve (leading 2 x 2 block [[2, 3], [3, 2]] is indefinite); a
    # singular Gram fails
    G = np.array([[2.0, 3.0, 0.5, 0.1], [3.0, 2.0, 0.2, 0.3], [0.5, 0.2, 1.5, 0.4],
                  [0.1, 0.3, 0.4, 1.2]])
    sigma = np.array([5.0, 4.0, 3.0, 2.0])
    Qm = rng.standard_normal((6, 4))
    m = ref.plssvd_weights(G, sigma, Qm)
    arrays = dict(G=G, sigma=sigma, Q=Qm, route=m["route"])
    for k, w in enumerate(m["W"]):
        arrays[f"W{k + 1}"] = w
    arrays["G_singular"] = np.array([[1.0, 1.0], [1.0, 1.0]])
    try:
        ref.plssvd_weights(arrays["G_singular"], sigma[:2], Qm[:, :2])
        raise SystemExit("the singular PLS-SVD case was solved")
    except RuntimeError:
        pass
    save("plssvd_pivoted", **arrays)


def special_cases(ref):
    """OPLS (the reference core's fit_opls / predict_opls for dense responses,
    the label filter + label SIMPLS for classes) and kernel PLS
    (fit_kernelpls / predict_kernelpls): the dense-SVD and sketch routes of
    the filter direction, radial-basis and polynomial kernels."""
    cases = [("opls_dense", dict(n=80, p=30, q=2, factors=6, seed=3, ncomp=4, north=2,
                                 oversample=4, power=2, rsvd_seed=7, scaling=2)),
             ("opls_sketch", dict(n=80, p=30, q=6, factors=6, seed=3, ncomp=4, north=2,
                                  oversample=2, power=1, rsvd_seed=7, scaling=1))]
    for name, c in cases:
        X, Y = latent_data(c["n"], c["p"], c["q"], factors=c["factors"], seed=c["seed"])
        Xt, _ = latent_data(20, c["p"], c["q"], factors=c["factors"], seed=c["seed"] + 1)
        m = ref.opls_fit(X, Y, ncomp=c["ncomp"], north=c["north"], scaling=c["scaling"],
                         oversample=c["oversample"], power=c["power"], seed=c["rsvd_seed"],
                         xtest=Xt, prefix=3)
        save(name, X=X, Y=Y, Xt=Xt, W=m["W"], P=m["P"], R=m["R"], Q=m["Q"], T=m["T"],
             mx=m["mx"], sx=m["sx"], my=m["my"], pred=m["pred"], prefix=np.int64(3),
             filtered=np.int64(m["filtered"]), completed=np.int64(m["completed"]),
             ncomp=np.int64(c["ncomp"]), north=np.int64(c["north"]),
             scaling=np.int64(c["scaling"]), oversample=np.int64(c["oversample"]),
             power=np.int64(c["power"]), seed=np.int64(c["rsvd_seed"]))
    X, lab = class_data(90, 25, 4, latent=4, seed=5)
    m = ref.opls_fit(X, labels=lab, classes=4, ncomp=3, north=2, oversample=1, power=1, seed=2)
    save("opls_labels", X=X, labels=np.asarray(lab, dtype=np.int32), classes=np.int64(4),
         W=m["W"], P=m["P"], R=m["R"], Q=m["Q"], T=m["T"], mx=m["mx"], sx=m["sx"], my=m["my"],
         filtered=np.int64(m["filtered"]), completed=np.int64(m["completed"]),
         ncomp=np.int64(3), north=np.int64(2), scaling=np.int64(2), oversample=np.int64(1),
         power=np.int64(1), seed=np.int64(2))
    X, Y = latent_data(70, 12, 2, factors=5, seed=11)
    Xt, _ = latent_data(15, 12, 2, factors=5, seed=12)
    for name, kernel, gamma, degree, coef0, scaling in (("kernelpls_rbf", 2, 0.05, 3, 1.0, 2),
                                                        ("kernelpls_poly", 3, 0.1, 2, 1.0, 1)):
        m = ref.kernelpls_fit(X, Y, ncomp=4, scaling=scaling, kernel=kernel, gamma=gamma,
                              degree=degree, coef0=coef0, oversample=4, power=2, seed=5,
                              xtest=Xt, prefix=3)
        save(name, X=X, Y=Y, Xt=Xt, R=m["R"], Q=m["Q"], T=m["T"], mx=m["mx"], sx=m["sx"],
             my=m["my"], colmeans=m["colmeans"], grand=np.float64(m["grand"]), pred=m["pred"],
             prefix=np.int64(3), completed=np.int64(m["completed"]), ncomp=np.int64(4),
             scaling=np.int64(scaling), kernel=np.int64(kernel), gamma=np.float64(gamma),
             degree=np.int64(degree), coef0=np.float64(coef0), oversample=np.int64(4),
             power=np.int64(2), seed=np.int64(5))


def reorth_cases(ref):
    """Dense SIMPLS with the re-orthogonalisation forced (the reference float
    core's routing: scores and weights against the earlier scores twice,
    simpls.hpp:161-200, 635-642, loadings twice): 30 components of a rank-8
    problem, where it and the single-pass routing part by about 5e-7 in B.
    Inputs are stored as their generator recipe and checksum."""
    n, p, q, factors, noise, seed, decay = 200, 300, 3, 8, 0.05, 7, 0.85
    X, Y = latent_data(n, p, q, factors=factors, noise=noise, seed=seed, decay=decay)
    ncomp, oversample, power, rseed = 30, 10, 2, 3
    m = ref.simpls_fit(X, Y, ncomp=ncomp, oversample=oversample, power=power, seed=rseed,
                       block=1, cache=0, reorth=1, batch_geometry=0)
    save("simpls_dense_reorth", R=m["R"], Q=m["Q"], V=m["V"], T=m["T"], mx=m["mx"], sx=m["sx"],
         my=m["my"], completed=np.int64(m["completed"]), ncomp=np.int64(ncomp),
         oversample=np.int64(oversample), power=np.int64(power), seed=np.int64(rseed),
         reorth=np.int64(1), recipe=np.asarray([n, p, q, factors, noise, seed, decay]),
         x_sha256=digest(X))


def main():
    build(reference=True)
    ref = Reference()
    if len(sys.argv) > 1 and sys.argv[1] == "reorth":
        reorth_cases(ref)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "audited":
        audited_cases(ref)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "operator":
        operator_cases(ref)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "edges":
        edge_cases(ref)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "special":
        special_cases(ref)
        return
    audited_cases(ref)
    operator_cases(ref)
    edge_cases(ref)
    special_cases(ref)
    reorth_cases(ref)

    # SIMPLS, dense responses, explicit width = q (direction independent of the sketch)
    X, Y = latent_data(64, 20, 3, factors=5, noise=0.1, seed=11)
    simpls_case(ref, "simpls_dense_small", X, Y, ncomp=5, seed=7)
    # SIMPLS with a sketch narrower than q (width 2 < q = 6): sketch-dependent
    X, Y = latent_data(300, 40, 6, factors=8, noise=0.1, seed=12)
    simpls_case(ref, "simpls_dense_sketch", X, Y, ncomp=8, oversample=1, power=2, seed=3)
    # SIMPLS classification (one-based labels). The centred one-hot
    # cross-product has rank K-1, so a sketch as wide as K makes the naive
    # test backend's Gram-Schmidt QR give up; widths below K keep the golden
    # run on the path that LAPACK-backed builds take.
    X, lab = class_data(400, 30, 4, latent=5, seed=13)
    simpls_case(ref, "simpls_labels_small", X, labels=lab, classes=4, ncomp=3, oversample=1,
                power=2, seed=5)
    # SIMPLS with a candidate block of 3 and batched candidate geometry
    X, lab = class_data(500, 60, 6, latent=6, seed=14)
    simpls_case(ref, "simpls_block3", X, labels=lab, classes=6, ncomp=5, oversample=1,
                power=2, seed=2, block=3, batch_geometry=1)
    # The plain SIMPLS regression workload: n=2000, p=1000, q=3, 10 components,
    # latent-factor data with seed 0, autoscaling; inputs regenerated from the recipe.
    X, Y = latent_data(2000, 1000, 3, factors=10, noise=0.1, seed=0)
    simpls_case(ref, "simpls_c1", X, Y, ncomp=10, seed=1, store_inputs=False,
                recipe=[2000, 1000, 3, 10, 0.1, 0])

    # PLS-SVD: randomized (width 6 < min(p,q) = 10) and dense (width >= q)
    X, Y = latent_data(120, 30, 10, factors=6, noise=0.1, seed=21)
    plssvd_case(ref, "plssvd_sketch", X, Y, ncomp=4, oversample=2, power=2, seed=17)
    X, Y = latent_data(100, 25, 4, factors=4, noise=0.1, seed=22)
    plssvd_case(ref, "plssvd_dense", X, Y, ncomp=3, oversample=32, power=5, seed=1)
    X, lab = class_data(300, 20, 5, latent=4, seed=23)
    # width 4 < 5 classes: the randomized route (the dense route of the naive test
    # backend rejects the rank-deficient centred one-hot cross-product)
    plssvd_case(ref, "plssvd_labels", X, labels=lab, classes=5, ncomp=3, oversample=1,
                power=2, seed=1)

    # Randomized SVD of a dense matrix
    A = np.random.default_rng(31).standard_normal((60, 40))
    r = ref.rsvd(A, 5, oversample=4, power=2, seed=11)
    save("rsvd_small", A=A, U=r["U"], s=r["s"], Vt=r["Vt"], k=np.int64(5),
         oversample=np.int64(4), power=np.int64(2), seed=np.int64(11))

    # LDA on scores
    rng = np.random.default_rng(41)
    T = rng.standard_normal((200, 4))
    lab = (np.arange(200) % 3 + 1).astype(np.int32)
    T[:, 0] += lab
    test = rng.standard_normal((50, 4))
    lda = ref.lda(T, lab, 3, 3, test)
    save("lda_small", T=T, labels=lab, test=test, disc=lda["disc"], predicted=lda["labels"],
         ridge=np.float64(lda["ridge"]), prefix=np.int64(3), classes=np.int64(3))


if __name__ == "__main__":
    main()
