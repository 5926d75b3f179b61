"""ctypes bindings for the FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker. The product path
(``fastpls_amd``) never imports it.

``Oracle`` wraps ``oracle/liboracle.so`` (the restatement in
``fastpls_oracle.cpp``); ``Reference`` wraps ``oracle/_ref/libfastpls_ref.so``
(the reference's own header core compiled by ``oracle/Makefile``), which is
present only where the reference checkout was available to build it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfastpls_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def build(reference: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if reference:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _labels(labels):
    if labels is None:
        return None
    return np.ascontiguousarray(np.asarray(labels, dtype=np.int32))



AUDIT_FIELDS = ("certified", "deterministic_fallback", "attempts", "effective_oversample",
                "effective_power", "effective_seed", "subspace_error", "singular_value_error",
                "triplet_residual", "omitted_direction_ratio")


def _lda_moments(fn, check, gram, class_sums, counts, cptr, n, prefix):
    g = _f64(gram); cs = _f64(class_sums)
    K, a = cs.shape
    lin = np.zeros((K, prefix), order="F"); con = np.zeros(K)
    ridge = C.c_double(0); rel = C.c_double(0)
    check(fn(_d(g), a, _d(cs), cptr(counts), K, C.c_longlong(int(n)), int(prefix), _d(lin),
             _d(con), C.byref(ridge), C.byref(rel)))
    return dict(linear=lin, constants=con, ridge=ridge.value, relative=rel.value)


def _plssvd_weights(fn, check, gram, sigma, Q):
    g = _f64(gram); s = np.ascontiguousarray(sigma, dtype=np.float64); q_ = _f64(Q)
    A, q = g.shape[0], q_.shape[0]
    W = np.zeros(A * A * q); route = np.zeros(A, dtype=np.int32)
    check(fn(_d(g), _d(s), _d(q_), A, q, _d(W), _i(route)))
    return dict(W=[W[(k - 1) * A * q:(k - 1) * A * q + k * q].reshape((k, q), order="F")
                   for k in range(1, A + 1)], route=route)


def _audited(lib_fn, check, a, k, oversample, power, seed, left_only):
    a = _f64(a); m, n = a.shape
    kk = min(k, m, n)
    U = np.zeros((m, kk), order="F"); s = np.zeros(kk); Vt = np.zeros((kk, n), order="F")
    got = C.c_int(0)
    audit = np.zeros(10)
    check(lib_fn(_d(a), m, n, k, oversample, power, C.c_uint(seed), int(bool(left_only)),
                 _d(U), _d(s), _d(Vt), C.byref(got), _d(audit)))
    g = got.value
    out = dict(U=U[:, :g], s=s[:g], Vt=None if left_only else np.asfortranarray(Vt[:g]))
    out["audit"] = dict(zip(AUDIT_FIELDS, audit.tolist()))
    return out


def _opls_call(fn, check, x, y, labels, classes, ncomp, north, scaling, oversample, power, seed,
               extra=(), with_reff=False, xtest=None, prefix=None):
    x = _f64(x); n, p = x.shape
    lab = _labels(labels)
    if lab is not None:
        q = int(classes); yy = None
    else:
        yy = _f64(y); q = yy.shape[1]
    A = max(1, min(ncomp, p, max(n - 1, 1)))
    W = np.zeros((p, max(north, 1)), order="F"); P = np.zeros((p, max(north, 1)), order="F")
    R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F"); T = np.zeros((n, A), order="F")
    Reff = np.zeros((p, A), order="F")
    mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
    filt = C.c_int(0); done = C.c_int(0)
    args = [_d(x), _d(yy), _i(lab), n, p, q, ncomp, north, scaling, oversample, power,
            C.c_uint(seed), *extra, _d(W), _d(P), C.byref(filt), _d(R), _d(Q), _d(T)]
    if with_reff:
        args.append(_d(Reff))
    args += [_d(mx), _d(sx), _d(my), C.byref(done)]
    pred = None
    if not with_reff:
        xt = _f64(xtest) if xtest is not None else None
        rows = 0 if xt is None else xt.shape[0]
        pred = np.zeros((max(rows, 1), q), order="F")
        args += [_d(xt), rows, int(prefix or A), _d(pred)]
        if rows == 0:
            pred = None
    check(fn(*args))
    f = filt.value
    out = dict(W=W[:, :f], P=P[:, :f], filtered=f, R=R, Q=Q, T=T, mx=mx, sx=sx, my=my,
               completed=done.value)
    if with_reff:
        out["Reff"] = Reff
    if pred is not None:
        out["pred"] = pred
    return out


def _kernelpls_call(fn, check, x, y, labels, classes, ncomp, scaling, kernel, gamma, degree,
                    coef0, oversample, power, seed, extra=(), xtest=None, prefix=None,
                    with_labels=True):
    x = _f64(x); n, p = x.shape
    lab = _labels(labels)
    if lab is not None:
        q = int(classes); yy = None
    else:
        yy = _f64(y); q = yy.shape[1]
    A = max(1, min(ncomp, max(n - 1, 1)))
    R = np.zeros((n, A), order="F"); Q = np.zeros((q, A), order="F"); T = np.zeros((n, A), order="F")
    mx = np.empty(p); sx = np.empty(p); my = np.empty(q); cm = np.empty(n)
    grand = C.c_double(0); done = C.c_int(0)
    xt = _f64(xtest) if xtest is not None else None
    rows = 0 if xt is None else xt.shape[0]
    pred = np.zeros((max(rows, 1), q), order="F")
    head = [_d(x), _d(yy)] + ([_i(lab)] if with_labels else [])
    check(fn(*head, n, p, q, ncomp, scaling, kernel, C.c_double(gamma), degree,
             C.c_double(coef0), oversample, power, C.c_uint(seed), *extra, _d(R), _d(Q), _d(T),
             _d(mx), _d(sx), _d(my), _d(cm), C.byref(grand), C.byref(done), _d(xt), rows,
             int(prefix or A), _d(pred)))
    out = dict(R=R, Q=Q, T=T, mx=mx, sx=sx, my=my, colmeans=cm, grand=grand.value,
               completed=done.value)
    if rows:
        out["pred"] = pred
    return out


class Oracle:
    """FP64 restatement of the reference algorithms (column-major)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.lib.orc_last_error.restype = C.c_char_p

    def _check(self, status):
        if status != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())

    def use_blas(self, threads: int | None = 1) -> str:
        """Route the oracle's products through numpy's bundled OpenBLAS dgemm
        (the reference CPU adapter's BLAS, for CPU-baseline timing) with
        `threads` BLAS threads; ``threads=None`` returns to the plain loops.
        Returns a description of the library bound."""
        if threads is None:
            self._check(self.lib.orc_use_blas(None, None, None, 0))
            return "plain loops"
        path = openblas_path()
        self._check(self.lib.orc_use_blas(path.encode(), b"scipy_dgemm_64_",
                                          b"scipy_openblas_set_num_threads64_", int(threads)))
        cfg = C.CDLL(path).scipy_openblas_get_config64_
        cfg.restype = C.c_char_p
        return f"{os.path.basename(path)} ({cfg().decode()}), {threads} thread(s)"

    def gaussian(self, seed: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.float64)
        self.lib.orc_gaussian(C.c_uint(seed), C.c_size_t(count), _d(out))
        return out

    def prepare(self, x, y=None, labels=None, classes=None, scaling=2):
        x = _f64(x)
        n, p = x.shape
        xs = np.empty((n, p), dtype=np.float64, order="F")
        mx = np.empty(p); sx = np.empty(p)
        if labels is not None:
            lab = _labels(labels)
            q = int(classes)
            S = np.empty((p, q), order="F"); my = np.empty(q)
            self._check(self.lib.orc_prepare_labels(_d(x), _i(lab), n, p, q, scaling,
                                                    _d(S), _d(mx), _d(sx), _d(my), _d(xs)))
        else:
            y = _f64(y)
            q = y.shape[1]
            S = np.empty((p, q), order="F"); my = np.empty(q)
            self._check(self.lib.orc_prepare_dense(_d(x), _d(y), n, p, q, scaling,
                                                   _d(S), _d(mx), _d(sx), _d(my), _d(xs)))
        return dict(S=S, mx=mx, sx=sx, my=my, xs=xs)

    def controls(self, n, p, q, A, classification=False):
        out = np.zeros(3, dtype=np.int32)
        self.lib.orc_simpls_controls(n, p, q, A, int(classification), _i(out))
        return dict(block=int(out[0]), cache=int(out[1]), batch_geometry=int(out[2]))

    def simpls_fit(self, x, y=None, labels=None, classes=None, ncomp=2, scaling=2,
                   oversample=32, power=5, seed=1, reorth=-1):
        x = _f64(x)
        n, p = x.shape
        lab = _labels(labels)
        if lab is not None:
            q = int(classes); yy = None
        else:
            yy = _f64(y); q = yy.shape[1]
        A = max(1, min(ncomp, p, max(n - 1, 1)))
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        V = np.zeros((p, A), order="F"); T = np.zeros((n, A), order="F")
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        done = C.c_int(0)
        self._check(self.lib.orc_simpls_fit(_d(x), _d(yy), _i(lab), n, p, q, ncomp, scaling,
                                            oversample, power, C.c_uint(seed), reorth,
                                            _d(R), _d(Q), _d(V), _d(T), _d(mx), _d(sx), _d(my),
                                            C.byref(done)))
        return dict(R=R, Q=Q, V=V, T=T, mx=mx, sx=sx, my=my, completed=done.value)

    def simpls_core(self, xs, S0, components, block=1, cache=0, reorth=0, batch_geometry=0,
                    oversample=32, power=5, seed=1):
        xs = _f64(xs); S0 = _f64(S0)
        n, p = xs.shape; q = S0.shape[1]
        A = max(1, min(components, p, max(n - 1, 1)))
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        V = np.zeros((p, A), order="F"); T = np.zeros((n, A), order="F")
        done = C.c_int(0)
        self._check(self.lib.orc_simpls_core(_d(xs), _d(S0), n, p, q, components, block, cache,
                                             reorth, batch_geometry, oversample, power,
                                             C.c_uint(seed), _d(R), _d(Q), _d(V), _d(T),
                                             C.byref(done)))
        return dict(R=R, Q=Q, V=V, T=T, completed=done.value)

    def rsvd(self, a, k, oversample=32, power=5, seed=1):
        a = _f64(a)
        m, n = a.shape
        kk = min(k, m, n)
        U = np.zeros((m, kk), order="F"); s = np.zeros(kk); Vt = np.zeros((kk, n), order="F")
        got = C.c_int(0)
        self._check(self.lib.orc_rsvd(_d(a), m, n, k, oversample, power, C.c_uint(seed),
                                      _d(U), _d(s), _d(Vt), C.byref(got)))
        g = got.value
        return dict(U=U[:, :g], s=s[:g], Vt=np.asfortranarray(Vt[:g]))

    def audited_rsvd(self, a, k, oversample=32, power=5, seed=1, left_only=False):
        """fastsvd: audited operator rSVD (audited_rsvd.hpp:298-414)."""
        return _audited(self.lib.orc_audited_rsvd, self._check, a, k, oversample, power, seed,
                        left_only)

    def fastcor(self, a, b=None, by_row=False, diagonal=False):
        """fastcor_values (src/r_api.cpp:4082-4142)."""
        a = _f64(a)
        bb = None if b is None else _f64(b)
        ua = a.shape[0] if by_row else a.shape[1]
        if bb is None:
            ub, diagonal = ua, False
        else:
            ub = bb.shape[0] if by_row else bb.shape[1]
        out = np.zeros(ua) if diagonal else np.zeros((ua, ub), order="F")
        self._check(self.lib.orc_fastcor(_d(a), a.shape[0], a.shape[1], _d(bb),
                                         0 if bb is None else bb.shape[0],
                                         0 if bb is None else bb.shape[1], int(bool(by_row)),
                                         int(bool(diagonal)), _d(out)))
        return out

    def plssvd_fit(self, x, y=None, labels=None, classes=None, ncomp=2, scaling=2,
                   oversample=32, power=5, seed=1):
        x = _f64(x)
        n, p = x.shape
        lab = _labels(labels)
        if lab is not None:
            q = int(classes); yy = None
        else:
            yy = _f64(y); q = yy.shape[1]
        A = ncomp
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        T = np.zeros((n, A), order="F"); s = np.zeros(A)
        W = np.zeros(A * A * q)
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        self._check(self.lib.orc_plssvd_fit(_d(x), _d(yy), _i(lab), n, p, q, A, scaling,
                                            oversample, power, C.c_uint(seed), _d(R), _d(Q),
                                            _d(T), _d(s), _d(W), _d(mx), _d(sx), _d(my)))
        weights = [W[k * A * q:k * A * q + (k + 1) * q].reshape((k + 1, q), order="F")
                   for k in range(A)]
        return dict(R=R, Q=Q, T=T, sigma=s, W=weights, mx=mx, sx=sx, my=my)

    def implicit_route(self, p, q):
        """R/main.R:3454-3463: the CPU path is matrix-free once p·q·8 > 32 MiB."""
        return bool(self.lib.orc_implicit_route(p, q))

    def simpls_operator_fit(self, x, y, ncomp=2, scaling=2, oversample=32, power=5, seed=1,
                            elem=8):
        """Matrix-free SIMPLS (simpls.hpp:768-1047) with the implicit route's
        controls for `elem`-byte entries (src/r_api.cpp:1466-1493)."""
        x = _f64(x); yy = _f64(y)
        n, p = x.shape; q = yy.shape[1]
        A = max(1, min(ncomp, p, max(n - 1, 1)))
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        V = np.zeros((p, A), order="F"); T = np.zeros((n, A), order="F")
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        done = C.c_int(0)
        route = np.zeros(1, dtype=np.int32)
        self._check(self.lib.orc_simpls_operator_fit(
            _d(x), _d(yy), n, p, q, ncomp, scaling, oversample, power, C.c_uint(seed), elem,
            _d(R), _d(Q), _d(V), _d(T), _d(mx), _d(sx), _d(my), C.byref(done), _i(route)))
        return dict(R=R, Q=Q, V=V, T=T, mx=mx, sx=sx, my=my, completed=done.value,
                    rank_one=int(route[0]))

    def simpls_operator_core(self, x, y, ncomp=2, scaling=2, oversample=32, power=5, seed=1,
                             block=1, cache=0, reorth=0, batch_geometry=0, rank_one=0):
        """Matrix-free SIMPLS with explicit controls."""
        x = _f64(x); yy = _f64(y)
        n, p = x.shape; q = yy.shape[1]
        A = ncomp
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        V = np.zeros((p, A), order="F"); T = np.zeros((n, A), order="F")
        done = C.c_int(0)
        self._check(self.lib.orc_simpls_operator_core(
            _d(x), _d(yy), n, p, q, A, scaling, oversample, power, C.c_uint(seed), block, cache,
            reorth, batch_geometry, rank_one, _d(R), _d(Q), _d(V), _d(T), C.byref(done)))
        return dict(R=R, Q=Q, V=V, T=T, completed=done.value)

    def plssvd_operator_fit(self, x, y, ncomp=2, scaling=2, oversample=32, power=5, seed=1):
        """Matrix-free PLS-SVD (plssvd.hpp:222-248)."""
        x = _f64(x); yy = _f64(y)
        n, p = x.shape; q = yy.shape[1]
        A = ncomp
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        T = np.zeros((n, A), order="F"); s = np.zeros(A)
        W = np.zeros(A * A * q)
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        self._check(self.lib.orc_plssvd_operator_fit(
            _d(x), _d(yy), n, p, q, A, scaling, oversample, power, C.c_uint(seed), _d(R),
            _d(Q), _d(T), _d(s), _d(W), _d(mx), _d(sx), _d(my)))
        weights = [W[k * A * q:k * A * q + (k + 1) * q].reshape((k + 1, q), order="F")
                   for k in range(A)]
        return dict(R=R, Q=Q, T=T, sigma=s, W=weights, mx=mx, sx=sx, my=my)

    def plssvd_core(self, xs, S, comps, oversample=32, power=5, seed=1):
        xs = _f64(xs); S = _f64(S)
        n, p = xs.shape; q = S.shape[1]
        comps = np.ascontiguousarray(np.asarray(comps, dtype=np.int32))
        A = int(comps.max())
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        T = np.zeros((n, A), order="F"); s = np.zeros(A)
        W = np.zeros((int(comps[-1]), q), order="F")
        self._check(self.lib.orc_plssvd_core(_d(xs), _d(S), n, p, q, _i(comps), len(comps),
                                             oversample, power, C.c_uint(seed), _d(R), _d(Q),
                                             _d(T), _d(s), _d(W)))
        return dict(R=R, Q=Q, T=T, sigma=s, W_last=W)

    def simpls_predict(self, model, x, prefix):
        x = _f64(x)
        rows, p = x.shape
        q = model["Q"].shape[0]; A = model["R"].shape[1]
        out = np.zeros((rows, q), order="F")
        self._check(self.lib.orc_simpls_predict(_d(x), rows, p, q, A, _d(_f64(model["R"])),
                                                _d(_f64(model["Q"])), _d(model["mx"]),
                                                _d(model["sx"]), _d(model["my"]), prefix,
                                                _d(out)))
        return out

    def lda(self, T, labels, classes, prefix, test_scores):
        T = _f64(T); n, a = T.shape
        ts = _f64(test_scores); rows = ts.shape[0]
        lab = _labels(labels)
        disc = np.zeros((rows, classes), order="F")
        pred = np.zeros(rows, dtype=np.int32)
        ridge = C.c_double(0)
        self._check(self.lib.orc_lda(_d(T), n, a, _i(lab), classes, prefix, _d(ts), rows,
                                     _d(disc), _i(pred), C.byref(ridge)))
        return dict(disc=disc, labels=pred, ridge=ridge.value)

    def lda_moments(self, gram, class_sums, counts, n, prefix):
        """LDA from score moments (lda.hpp:226-346)."""
        return _lda_moments(self.lib.orc_lda_moments, self._check, gram, class_sums,
                            np.ascontiguousarray(counts, dtype=np.int32), _i, n, prefix)

    def opls_fit(self, x, y=None, labels=None, classes=None, ncomp=2, north=1, scaling=2,
                 oversample=32, power=5, seed=1):
        """OPLS (orthogonal filter + SIMPLS on the filtered predictors) with
        R_eff, the weights that take standardised rows to the scores."""
        return _opls_call(self.lib.orc_opls_fit, self._check, x, y, labels, classes, ncomp,
                          north, scaling, oversample, power, seed, with_reff=True)

    def kernelpls_fit(self, x, y=None, labels=None, classes=None, ncomp=2, scaling=2, kernel=2,
                      gamma=1.0, degree=3, coef0=1.0, oversample=32, power=5, seed=1,
                      xtest=None, prefix=None):
        return _kernelpls_call(self.lib.orc_kernelpls_fit, self._check, x, y, labels, classes,
                               ncomp, scaling, kernel, gamma, degree, coef0, oversample, power,
                               seed, xtest=xtest, prefix=prefix)

    def plssvd_weights(self, gram, sigma, Q):
        """PLS-SVD prefix weights (plssvd.hpp:138-164) and the solve route per prefix."""
        return _plssvd_weights(self.lib.orc_plssvd_weights, self._check, gram, sigma, Q)

    def topk(self, values, top):
        v = _f64(values); rows, classes = v.shape
        out = np.zeros((rows, top), dtype=np.int32, order="F")
        self.lib.orc_topk(_d(v), rows, classes, top, out.ctypes.data_as(_ip))
        return out

    def response_sums(self, pred, train_mean, observed=None, labels=None):
        pred = _f64(pred); n, q = pred.shape
        out = np.zeros(3 * q)
        self._check(self.lib.orc_response_sums(_d(pred), _d(_f64(observed)) if observed is not None else None,
                                               _i(_labels(labels)), _d(np.asarray(train_mean, dtype=np.float64)),
                                               n, q, _d(out)))
        return out.reshape((3, q), order="F")

    def variance(self, xs, T):
        xs = _f64(xs); T = _f64(T)
        n, p = xs.shape; a = T.shape[1]
        P = np.zeros((p, a), order="F"); ss = np.zeros(a + 1)
        self._check(self.lib.orc_variance(_d(xs), _d(T), n, p, a, _d(P), _d(ss)))
        return dict(P=P, ss=ss)

    def gemm(self, left, right):
        left = _f64(left); right = _f64(right)
        out = np.zeros((left.shape[0], right.shape[1]), order="F")
        self.lib.orc_gemm(_d(left), _d(right), left.shape[0], left.shape[1], right.shape[1], _d(out))
        return out

    def eigen(self, a):
        a = _f64(a).copy(order="F"); n = a.shape[0]
        vals = np.zeros(n)
        self._check(self.lib.orc_symmetric_eigen(_d(a), n, _d(vals)))
        return vals, a

    def qr(self, a):
        a = _f64(a); m, n = a.shape
        q = np.zeros((m, min(m, n)), order="F")
        self._check(self.lib.orc_qr(_d(a), m, n, _d(q)))
        return q


class Reference:
    """The reference's own header core with its naive test backend."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.lib.ref_last_error.restype = C.c_char_p

    def _check(self, status):
        if status != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def block_size(self, remaining, p, q, classification, n, maximum):
        f = self.lib.ref_block_size
        f.restype = C.c_longlong
        f.argtypes = [C.c_longlong] * 3 + [C.c_int] + [C.c_longlong] * 2
        return int(f(remaining, p, q, int(classification), n, maximum))

    def simpls_fit(self, x, y=None, labels=None, classes=None, ncomp=2, scaling=2,
                   oversample=32, power=5, seed=1, block=1, cache=0, reorth=0,
                   batch_geometry=0):
        x = _f64(x); n, p = x.shape
        lab = _labels(labels)
        if lab is not None:
            q = int(classes); yy = None
        else:
            yy = _f64(y); q = yy.shape[1]
        A = ncomp
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        V = np.zeros((p, A), order="F"); T = np.zeros((n, A), order="F")
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        done = C.c_int(0)
        self._check(self.lib.ref_simpls_fit(_d(x), _d(yy), _i(lab), n, p, q, A, scaling,
                                            oversample, power, C.c_uint(seed), block, cache,
                                            reorth, batch_geometry, _d(R), _d(Q), _d(V), _d(T),
                                            _d(mx), _d(sx), _d(my), C.byref(done)))
        return dict(R=R, Q=Q, V=V, T=T, mx=mx, sx=sx, my=my, completed=done.value)

    def plssvd_fit(self, x, y=None, labels=None, classes=None, ncomp=2, scaling=2,
                   oversample=32, power=5, seed=1):
        x = _f64(x); n, p = x.shape
        lab = _labels(labels)
        if lab is not None:
            q = int(classes); yy = None
        else:
            yy = _f64(y); q = yy.shape[1]
        A = ncomp
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        T = np.zeros((n, A), order="F"); s = np.zeros(A)
        W = np.zeros(A * A * q)
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        self._check(self.lib.ref_plssvd_fit(_d(x), _d(yy), _i(lab), n, p, q, A, scaling,
                                            oversample, power, C.c_uint(seed), _d(R), _d(Q),
                                            _d(T), _d(s), _d(W), _d(mx), _d(sx), _d(my)))
        weights = [W[k * A * q:k * A * q + (k + 1) * q].reshape((k + 1, q), order="F")
                   for k in range(A)]
        return dict(R=R, Q=Q, T=T, sigma=s, W=weights, mx=mx, sx=sx, my=my)

    def simpls_operator_fit(self, x, y, ncomp=2, scaling=2, oversample=32, power=5, seed=1,
                            block=1, cache=0, reorth=0, batch_geometry=0, rank_one=0):
        x = _f64(x); yy = _f64(y)
        n, p = x.shape; q = yy.shape[1]
        A = ncomp
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        V = np.zeros((p, A), order="F"); T = np.zeros((n, A), order="F")
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        done = C.c_int(0)
        self._check(self.lib.ref_simpls_operator_fit(
            _d(x), _d(yy), n, p, q, A, scaling, oversample, power, C.c_uint(seed), block, cache,
            reorth, batch_geometry, rank_one, _d(R), _d(Q), _d(V), _d(T), _d(mx), _d(sx),
            _d(my), C.byref(done)))
        return dict(R=R, Q=Q, V=V, T=T, mx=mx, sx=sx, my=my, completed=done.value)

    def plssvd_operator_fit(self, x, y, ncomp=2, scaling=2, oversample=32, power=5, seed=1):
        x = _f64(x); yy = _f64(y)
        n, p = x.shape; q = yy.shape[1]
        A = ncomp
        R = np.zeros((p, A), order="F"); Q = np.zeros((q, A), order="F")
        T = np.zeros((n, A), order="F"); s = np.zeros(A)
        W = np.zeros(A * A * q)
        mx = np.empty(p); sx = np.empty(p); my = np.empty(q)
        self._check(self.lib.ref_plssvd_operator_fit(
            _d(x), _d(yy), n, p, q, A, scaling, oversample, power, C.c_uint(seed), _d(R), _d(Q),
            _d(T), _d(s), _d(W), _d(mx), _d(sx), _d(my)))
        weights = [W[k * A * q:k * A * q + (k + 1) * q].reshape((k + 1, q), order="F")
                   for k in range(A)]
        return dict(R=R, Q=Q, T=T, sigma=s, W=weights, mx=mx, sx=sx, my=my)

    def rsvd(self, a, k, oversample=32, power=5, seed=1):
        a = _f64(a); m, n = a.shape
        kk = min(k, m, n)
        U = np.zeros((m, kk), order="F"); s = np.zeros(kk); Vt = np.zeros((kk, n), order="F")
        got = C.c_int(0)
        self._check(self.lib.ref_rsvd(_d(a), m, n, k, oversample, power, C.c_uint(seed),
                                      _d(U), _d(s), _d(Vt), C.byref(got)))
        g = got.value
        return dict(U=U[:, :g], s=s[:g], Vt=np.asfortranarray(Vt[:g]))

    def audited_rsvd(self, a, k, oversample=32, power=5, seed=1, left_only=False):
        return _audited(self.lib.ref_audited_rsvd, self._check, a, k, oversample, power, seed,
                        left_only)

    def lda_moments(self, gram, class_sums, counts, n, prefix):
        return _lda_moments(self.lib.ref_lda_moments, self._check, gram, class_sums,
                            np.ascontiguousarray(counts, dtype=np.float64), _d, n, prefix)

    def opls_fit(self, x, y=None, labels=None, classes=None, ncomp=2, north=1, scaling=2,
                 oversample=32, power=5, seed=1, block=1, cache=0, reorth=0, batch_geometry=0,
                 xtest=None, prefix=None):
        return _opls_call(self.lib.ref_opls_fit, self._check, x, y, labels, classes, ncomp,
                          north, scaling, oversample, power, seed,
                          extra=(block, cache, reorth, batch_geometry), xtest=xtest,
                          prefix=prefix)

    def kernelpls_fit(self, x, y, ncomp=2, scaling=2, kernel=2, gamma=1.0, degree=3, coef0=1.0,
                      oversample=32, power=5, seed=1, block=1, cache=0, reorth=0,
                      batch_geometry=0, xtest=None, prefix=None):
        return _kernelpls_call(self.lib.ref_kernelpls_fit, self._check, x, y, None, None, ncomp,
                               scaling, kernel, gamma, degree, coef0, oversample, power, seed,
                               extra=(block, cache, reorth, batch_geometry), xtest=xtest,
                               prefix=prefix, with_labels=False)

    def plssvd_weights(self, gram, sigma, Q):
        return _plssvd_weights(self.lib.ref_plssvd_weights, self._check, gram, sigma, Q)

    def lda(self, T, labels, classes, prefix, test_scores):
        T = _f64(T); n, a = T.shape
        ts = _f64(test_scores)[:, :prefix].copy(order="F"); rows = ts.shape[0]
        lab = _labels(labels)
        disc = np.zeros((rows, classes), order="F")
        pred = np.zeros(rows, dtype=np.int32)
        ridge = C.c_double(0)
        self._check(self.lib.ref_lda(_d(T), n, a, _i(lab), classes, prefix, _d(ts), rows,
                                     _d(disc), _i(pred), C.byref(ridge)))
        return dict(disc=disc, labels=pred, ridge=ridge.value)


def openblas_path() -> str:
    """numpy's bundled OpenBLAS (scipy_openblas64_, 64-bit integer interface)."""
    import glob
    libs = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), "numpy.libs")
    found = sorted(glob.glob(os.path.join(libs, "libscipy_openblas64_*.so")))
    if not found:
        raise RuntimeError(f"no bundled OpenBLAS under {libs}")
    return found[0]


def latent_data(n, p, q, factors=10, noise=0.1, seed=0, decay=None, dtype=np.float64):
    """Latent-factor regression data: X = T Pᵀ + noise·E, Y = T Qᵀ + noise·F."""
    rng = np.random.default_rng(seed)
    T = rng.standard_normal((n, factors))
    if decay is not None:
        T *= decay ** np.arange(factors)
    P = rng.standard_normal((p, factors))
    Qm = rng.standard_normal((q, factors))
    X = T @ P.T + noise * rng.standard_normal((n, p))
    Y = T @ Qm.T + noise * rng.standard_normal((n, q))
    return np.asfortranarray(X.astype(dtype)), np.asfortranarray(Y.astype(dtype))


def class_data(n, p, K, latent=15, seed=2, spread=0.5, dtype=np.float64):
    """Balanced classes whose means live in a latent subspace."""
    rng = np.random.default_rng(seed)
    labels = (np.arange(n) % K) + 1
    rng.shuffle(labels)
    means = spread * rng.standard_normal((K, latent))
    L = rng.standard_normal((p, latent))
    Z = means[labels - 1]
    X = Z @ L.T + rng.standard_normal((n, p))
    return np.asfortranarray(X.astype(dtype)), labels.astype(np.int32)
