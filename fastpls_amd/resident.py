"""Python binding of the HIP C ABI (include/fastpls_hip.h).

Mirrors the reference's resident-model glue (src/cuda_resident_r.cpp and
R/resident_cuda.R): a model handle created from predictors plus responses or
one-based labels, with prediction / classification paths over component
prefixes, response sums, projection, export of the fitted factors, controls
and compaction. Arrays are numpy, column-major (Fortran order) like the
reference's R matrices. There is no CPU fallback: if the HIP library is
missing or no device is usable, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfastpls_hip.so")

SIMPLS = 0
PLSSVD = 1
OPLS = 2       # handle tags of the special models (not method codes of the C ABI)
KERNELPLS = 3

FIELDS = {"R": 0, "Q": 1, "T": 2, "mX": 3, "vX": 4, "mY": 5, "P": 6, "variance": 7}

# name -> (restype, argtypes)
_vp, _cvp, _ip, _dp, _cp = C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p
_sz, _ull, _ll = C.c_size_t, C.c_ulonglong, C.c_longlong
_CV_REGRESSION = (C.c_int, [_cvp, _cvp, _ip, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                            C.c_int, C.c_int, C.c_int, C.c_int, _ull, C.c_int, _dp, _ip, _dp,
                            _dp, _dp, _dp, _cp, _sz])
_CV_CLASSIFICATION = (C.c_int, [_cvp, _ip, _ip, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                C.c_int, C.c_int, C.c_int, C.c_int, _ull, C.c_int, C.c_int, _ip,
                                _vp, _ip, _dp, _cp, _sz])

SIGNATURES = {
    "fastpls_hip_server_create": (_vp, [C.c_double, C.c_int, C.c_int, C.c_int, _cp, _sz]),
    "fastpls_hip_server_submit_fit": (_ll, [_vp, C.c_int, _cvp, _cvp, _ip, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, _ull, _cp, _sz]),
    "fastpls_hip_server_submit_predict": (_ll, [_vp, _ll, _cvp, _ll, C.c_int, C.c_int, C.c_int,
                                                _vp, _cp, _sz]),
    "fastpls_hip_server_poll": (C.c_int, [_vp, _ll, C.c_int, _ip, _dp, _cp, _sz]),
    "fastpls_hip_server_release": (C.c_int, [_vp, _ll, _cp, _sz]),
    "fastpls_hip_server_export": (C.c_int, [_vp, _ll, C.c_int, _vp, _sz, _cp, _sz]),
    "fastpls_hip_server_stats": (C.c_int, [_vp, _dp, C.c_int, _cp, _sz]),
    "fastpls_hip_server_destroy": (None, [_vp]),
    "fastpls_hip_device_info": (C.c_int, [_cp, _sz]),
    "fastpls_hip_gemm": (C.c_int, [_cvp, _cvp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _cp, _sz]),
    "fastpls_hip_simpls_create": (_vp, [_cvp, _cvp, _ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ull, _cp, _sz]),
    "fastpls_hip_plssvd_create": (_vp, [_cvp, _cvp, _ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ull, _cp, _sz]),
    "fastpls_hip_opls_create": (_vp, [_cvp, _cvp, _ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int, _ull, C.c_int, _cp, _sz]),
    "fastpls_hip_kernelpls_create": (_vp, [_cvp, _cvp, _ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ull,
                                           C.c_int, C.c_double, C.c_int, C.c_double, _cp, _sz]),
    "fastpls_hip_predict_path": (C.c_int, [_vp, _cvp, C.c_int, _ip, C.c_int, C.c_int, _vp, _cp, _sz]),
    "fastpls_hip_predict_noalloc": (C.c_int, [_vp, _cvp, C.c_int, C.c_int, _vp, _cp, _sz]),
    "fastpls_hip_destroy": (None, [_vp]),
    "fastpls_hip_export": (C.c_int, [_vp, C.c_int, _vp, _sz, _cp, _sz]),
    "fastpls_hip_classify_path": (C.c_int, [_vp, _cvp, C.c_int, _ip, C.c_int, C.c_int, C.c_int,
                                            _ip, _cp, _sz]),
    "fastpls_hip_classify_response_path": (C.c_int, [_vp, _cvp, C.c_int, _ip, C.c_int, C.c_int,
                                                     C.c_int, _ip, _vp, _cp, _sz]),
    "fastpls_hip_response_sums": (C.c_int, [_vp, _cvp, _cvp, _ip, C.c_int, C.c_int, _vp, _cp, _sz]),
    "fastpls_hip_project": (C.c_int, [_vp, _cvp, C.c_int, C.c_int, _vp, _cp, _sz]),
    "fastpls_hip_controls": (C.c_int, [_vp, _ip, _ip, _ip, _ip, _ip, _ip, _cp, _sz]),
    "fastpls_hip_compact": (C.c_int, [_vp, C.c_int, _cp, _sz]),
    "fastpls_hip_fit_device": (_vp, [C.c_int, _cvp, _cvp, _ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_int, _ull, _cp, _sz]),
    "fastpls_hip_last_fit_seconds": (C.c_int, [_vp, _dp, _cp, _sz]),
    "fastpls_hip_fit_phases": (C.c_int, [_vp, _dp, C.c_int, _cp, _sz]),
    "fastpls_hip_device_alloc": (C.c_int, [_sz, C.POINTER(C.c_void_p), _cp, _sz]),
    "fastpls_hip_device_free": (C.c_int, [_vp, _cp, _sz]),
    "fastpls_hip_memcpy": (C.c_int, [_vp, _cvp, _sz, C.c_int, _cp, _sz]),
    "fastpls_hip_latent_data": (C.c_int, [_vp, _vp, C.c_int, _ll, C.c_int, C.c_int, C.c_int,
                                          C.c_double, C.c_double, _ull, C.c_int, _cp, _sz]),
    "fastpls_hip_simpls_batch_create": (_vp, [C.c_int, C.c_int, _vp, _ip, _ip, C.c_int, _ip, _vp,
                                              _vp, _vp, _ip, _ip, C.POINTER(_ull), C.c_int,
                                              C.c_int, C.c_int, C.c_int, _cp, _sz]),
    "fastpls_hip_batch_status": (C.c_int, [_vp, _ip, _ip, _dp, _cp, _sz]),
    "fastpls_hip_batch_export": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _sz, _cp, _sz]),
    "fastpls_hip_batch_phases": (C.c_int, [_vp, _vp, C.c_int, _cp, _sz]),
    "fastpls_hip_batch_export64": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _sz, _cp, _sz]),
    "fastpls_hip_batch_rerun": (C.c_int, [_vp, _cp, _sz]),
    "fastpls_hip_batch_autotune": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _cp, _sz]),
    "fastpls_hip_batch_splits": (C.c_int, [_vp, C.POINTER(C.c_longlong), _cp, _sz]),
    "fastpls_hip_batch_lanes": (C.c_int, [_vp, C.POINTER(C.c_int), _cp, _sz]),
    "fastpls_hip_batch_set_splits": (C.c_int, [C.POINTER(C.c_longlong), _cp, _sz]),
    "fastpls_hip_batch_destroy": (None, [_vp]),
    "fastpls_hip_simpls_cv_regression": _CV_REGRESSION,
    "fastpls_hip_plssvd_cv_regression": _CV_REGRESSION,
    "fastpls_hip_simpls_cv_classification": _CV_CLASSIFICATION,
    "fastpls_hip_plssvd_cv_classification": _CV_CLASSIFICATION,
    "fastpls_hip_rsvd": (C.c_int, [_cvp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_uint, C.c_int, _dp, _dp, _dp, _ip, _dp, _cp, _sz]),
    "fastpls_hip_lda_moments": (C.c_int, [_dp, C.c_int, _dp, _ip, C.c_int, _ll, C.c_int, _dp,
                                          _dp, _dp, _cp, _sz]),
    "fastpls_hip_plssvd_weights": (C.c_int, [_dp, _dp, _dp, C.c_int, C.c_int, _dp, _cp, _sz]),
    "fastpls_hip_core_workspace_create_f32": (_vp, [_cp, _sz]),
    "fastpls_hip_core_workspace_destroy_f32": (None, [_vp]),
    "fastpls_hip_core_gemm_into_f32": (C.c_int, [_vp, _cvp, _sz, _sz, _sz, _cvp, _sz, _sz, _sz,
                                                 C.c_int, C.c_int, _vp, _sz, _sz, _sz, _cp, _sz]),
    "fastpls_hip_core_self_gram_into_f32": (C.c_int, [_vp, _cvp, _sz, _sz, _sz, C.c_int, _vp,
                                                      _sz, _sz, _sz, C.c_int, _cp, _sz]),
    "fastpls_hip_core_gemm_f32": (C.c_int, [_cvp, _sz, _sz, _sz, _cvp, _sz, _sz, _sz, C.c_int,
                                            C.c_int, _vp, _cp, _sz]),
    "fastpls_hip_fastcor": (C.c_int, [_cvp, C.c_int, C.c_int, _cvp, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, _dp, _cp, _sz]),
}

_LIB = None


class HipError(RuntimeError):
    """Raised for every failure reported by the HIP library."""


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the in-tree HIP library (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise HipError(f"HIP library not built: {path} (run python -m fastpls_amd.build)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _err():
    return C.create_string_buffer(1024)


def _raise(buf):
    raise HipError(buf.value.decode(errors="replace") or "unknown fastPLS HIP failure")


def _dtype(precision: int):
    if precision == 32:
        return np.float32
    if precision == 64:
        return np.float64
    raise ValueError("precision must be 32 or 64")


def _fmat(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def device_info() -> str:
    lib = load_library()
    buf = C.create_string_buffer(256)
    if lib.fastpls_hip_device_info(buf, 256) != 0:
        _raise(buf)
    return buf.value.decode()


def gemm(left, right, precision: int = 64):
    """fastpls_cuda_gemm counterpart: left @ right on the device."""
    lib = load_library()
    dt = _dtype(precision)
    a = _fmat(left, dt)
    b = _fmat(right, dt)
    if a.shape[1] != b.shape[0]:
        raise ValueError("inner dimensions differ")
    out = np.empty((a.shape[0], b.shape[1]), dtype=dt, order="F")
    e = _err()
    if lib.fastpls_hip_gemm(_ptr(a), _ptr(b), precision, a.shape[0], a.shape[1], b.shape[1],
                            _ptr(out), e, 1024) != 0:
        _raise(e)
    return out


class ResidentModel:
    """A fitted model resident in HBM (handle of the C ABI)."""

    def __init__(self, handle, method, precision, n, p, q, components, classification,
                 feature_p=None):
        self._h = handle
        self.method = method
        self.precision = precision
        self.dtype = _dtype(precision)
        self.n, self.p, self.q = n, p, q
        # columns of the weights R and loadings P: p, or n for kernel PLS
        self.feature_p = p if feature_p is None else feature_p
        self.components = components
        self.classification = classification

    # ----------------------------------------------------------- creation
    @staticmethod
    def _inputs(x, y, labels, classes, dt):
        xa = _fmat(x, dt)
        n, p = xa.shape
        if (y is None) == (labels is None):
            raise ValueError("provide responses or labels, not both")
        if labels is not None:
            lab = np.ascontiguousarray(np.asarray(labels, dtype=np.int32))
            q = int(classes if classes is not None else lab.max())
            ya = None
        else:
            ya = _fmat(y, dt)
            if ya.ndim == 1:
                ya = ya.reshape((-1, 1), order="F")
            q = ya.shape[1]
            lab = None
        return xa, ya, lab, n, p, q

    @classmethod
    def fit(cls, x, y=None, labels=None, classes=None, method=SIMPLS, precision=64,
            components=2, scaling=2, oversample=32, power=5, seed=1, retain_scores=True):
        """Fit SIMPLS (method=0) or PLS-SVD (method=1) from host arrays."""
        lib = load_library()
        xa, ya, lab, n, p, q = cls._inputs(x, y, labels, classes, _dtype(precision))
        e = _err()
        fn = lib.fastpls_hip_simpls_create if method == SIMPLS else lib.fastpls_hip_plssvd_create
        h = fn(_ptr(xa), _ptr(ya), _iptr(lab), precision, n, p, q, components, scaling,
               oversample, power, int(bool(retain_scores)), seed, e, 1024)
        if not h:
            _raise(e)
        return cls(h, method, precision, n, p, q, components, lab is not None)

    @classmethod
    def fit_opls(cls, x, y=None, labels=None, classes=None, precision=64, components=2, north=1,
                 scaling=2, oversample=32, power=5, seed=1, retain_scores=True):
        """Resident OPLS (fastpls_hip_opls_create): `north` orthogonal filter
        components, then SIMPLS on the filtered predictors."""
        lib = load_library()
        xa, ya, lab, n, p, q = cls._inputs(x, y, labels, classes, _dtype(precision))
        e = _err()
        h = lib.fastpls_hip_opls_create(_ptr(xa), _ptr(ya), _iptr(lab), precision, n, p, q,
                                        components, scaling, oversample, power,
                                        int(bool(retain_scores)), seed, north, e, 1024)
        if not h:
            _raise(e)
        return cls(h, OPLS, precision, n, p, q, components, lab is not None)

    @classmethod
    def fit_kernelpls(cls, x, y=None, labels=None, classes=None, precision=64, components=2,
                      kernel=2, gamma=1.0, degree=3, coef0=1.0, scaling=2, oversample=32,
                      power=5, seed=1, retain_scores=True):
        """Resident nonlinear kernel PLS (fastpls_hip_kernelpls_create):
        kernel 2 radial basis, 3 polynomial; SIMPLS on the centred kernel."""
        lib = load_library()
        xa, ya, lab, n, p, q = cls._inputs(x, y, labels, classes, _dtype(precision))
        e = _err()
        h = lib.fastpls_hip_kernelpls_create(_ptr(xa), _ptr(ya), _iptr(lab), precision, n, p, q,
                                             components, scaling, oversample, power,
                                             int(bool(retain_scores)), seed, kernel,
                                             float(gamma), degree, float(coef0), e, 1024)
        if not h:
            _raise(e)
        return cls(h, KERNELPLS, precision, n, p, q, components, lab is not None, feature_p=n)

    @classmethod
    def fit_device(cls, x_ptr, y_ptr, labels_ptr, n, p, q, method=SIMPLS, precision=32,
                   components=2, scaling=2, oversample=32, power=5, seed=1):
        """Fit from device pointers (integers) that stay owned by the caller."""
        lib = load_library()
        e = _err()
        lab = C.cast(C.c_void_p(labels_ptr), _ip) if labels_ptr else None
        h = lib.fastpls_hip_fit_device(method, C.c_void_p(x_ptr),
                                       C.c_void_p(y_ptr) if y_ptr else None, lab, precision,
                                       n, p, q, components, scaling, oversample, power, seed,
                                       e, 1024)
        if not h:
            _raise(e)
        return cls(h, method, precision, n, p, q, components, labels_ptr is not None)

    def close(self):
        if self._h:
            load_library().fastpls_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -------------------------------------------------------------- calls
    def _x(self, x):
        xa = _fmat(x, self.dtype)
        if xa.ndim != 2 or xa.shape[1] != self.p:
            raise ValueError("new data must have p columns")
        return xa

    @staticmethod
    def _prefixes(prefixes):
        return np.ascontiguousarray(np.atleast_1d(np.asarray(prefixes, dtype=np.int32)))

    def predict_path(self, x, prefixes, lda=False):
        xa = self._x(x)
        pf = self._prefixes(prefixes)
        rows = xa.shape[0]
        out = np.empty((rows, self.q, len(pf)), dtype=self.dtype, order="F")
        e = _err()
        if load_library().fastpls_hip_predict_path(self._h, _ptr(xa), rows, _iptr(pf), len(pf),
                                                   int(bool(lda)), _ptr(out), e, 1024) != 0:
            _raise(e)
        return out

    def predict(self, x, prefix, lda=False):
        return self.predict_path(x, [prefix], lda)[:, :, 0]

    def predict_noalloc(self, x_device, rows, prefix, out_device):
        """Responses of one prefix for device rows (pointer) into a device
        output (pointer, rows x q, model precision), allocating nothing
        (fastpls_hip_predict_noalloc)."""
        e = _err()
        if load_library().fastpls_hip_predict_noalloc(self._h, C.c_void_p(x_device), int(rows),
                                                      int(prefix), C.c_void_p(out_device), e,
                                                      1024) != 0:
            _raise(e)

    def classify_path(self, x, prefixes, lda=False, top=1):
        xa = self._x(x)
        pf = self._prefixes(prefixes)
        rows = xa.shape[0]
        out = np.empty((rows, top, len(pf)), dtype=np.int32, order="F")
        e = _err()
        if load_library().fastpls_hip_classify_path(self._h, _ptr(xa), rows, _iptr(pf), len(pf),
                                                    int(bool(lda)), top, _iptr(out), e, 1024) != 0:
            _raise(e)
        return out

    def classify_response_path(self, x, prefixes, lda=False, top=1):
        xa = self._x(x)
        pf = self._prefixes(prefixes)
        rows = xa.shape[0]
        labels = np.empty((rows, top, len(pf)), dtype=np.int32, order="F")
        resp = np.empty((rows, self.q, len(pf)), dtype=self.dtype, order="F")
        e = _err()
        if load_library().fastpls_hip_classify_response_path(
                self._h, _ptr(xa), rows, _iptr(pf), len(pf), int(bool(lda)), top, _iptr(labels),
                _ptr(resp), e, 1024) != 0:
            _raise(e)
        return labels, resp

    def response_sums(self, x, prefix, y=None, labels=None):
        xa = self._x(x)
        rows = xa.shape[0]
        ya = None if y is None else _fmat(y, self.dtype)
        lab = None if labels is None else np.ascontiguousarray(np.asarray(labels, dtype=np.int32))
        out = np.empty((3, self.q), dtype=self.dtype, order="F")
        e = _err()
        if load_library().fastpls_hip_response_sums(self._h, _ptr(xa), _ptr(ya), _iptr(lab), rows,
                                                     prefix, _ptr(out), e, 1024) != 0:
            _raise(e)
        return out

    def project(self, x, prefix):
        xa = self._x(x)
        rows = xa.shape[0]
        out = np.empty((rows, prefix), dtype=self.dtype, order="F")
        e = _err()
        if load_library().fastpls_hip_project(self._h, _ptr(xa), rows, prefix, _ptr(out), e,
                                              1024) != 0:
            _raise(e)
        return out

    def export(self, field):
        f = FIELDS[field] if isinstance(field, str) else int(field)
        a = self.components
        fp = self.feature_p
        shapes = {0: (fp, a), 1: (self.q, a), 2: (self.n, a), 3: (self.p,), 4: (self.p,),
                  5: (self.q,), 6: (fp, a), 7: (a + 1,)}
        shape = shapes[f]
        out = np.empty(shape, dtype=self.dtype, order="F")
        e = _err()
        if load_library().fastpls_hip_export(self._h, f, _ptr(out), out.size, e, 1024) != 0:
            _raise(e)
        return out

    def controls(self):
        vals = [C.c_int(0) for _ in range(6)]
        e = _err()
        if load_library().fastpls_hip_controls(self._h, *[C.byref(v) for v in vals], e, 1024) != 0:
            _raise(e)
        keys = ["oversample", "power", "block", "block_limit", "implicit_operator",
                "predictor_crossprod_cache"]
        return {k: v.value for k, v in zip(keys, vals)}

    def compact(self, prepare_lda=False):
        e = _err()
        if load_library().fastpls_hip_compact(self._h, int(bool(prepare_lda)), e, 1024) != 0:
            _raise(e)

    def last_fit_seconds(self) -> float:
        v = C.c_double(0)
        e = _err()
        if load_library().fastpls_hip_last_fit_seconds(self._h, C.byref(v), e, 1024) != 0:
            _raise(e)
        return v.value

    def fit_phases(self):
        out = (C.c_double * 6)()
        e = _err()
        if load_library().fastpls_hip_fit_phases(self._h, out, 6, e, 1024) != 0:
            _raise(e)
        return {"fit": out[0], "crossprod_pass": out[1], "score_pass": out[2],
                "x_bytes": out[3], "projection": out[4], "projection_bytes": out[5]}

    def coefficients(self, prefix):
        """B for one prefix (p x q) in standardized units: R_k Q_kᵀ (SIMPLS)."""
        R = self.export("R").astype(np.float64)
        Q = self.export("Q").astype(np.float64)
        return R[:, :prefix] @ Q[:, :prefix].T


class DeviceBuffer:
    """Device memory owned through the library (no torch in the process)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        ptr = C.c_void_p(0)
        e = _err()
        if load_library().fastpls_hip_device_alloc(self.nbytes, C.byref(ptr), e, 1024) != 0:
            _raise(e)
        self.ptr = ptr.value

    @classmethod
    def from_array(cls, a):
        a = np.asfortranarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a):
        a = np.asfortranarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        e = _err()
        if load_library().fastpls_hip_memcpy(C.c_void_p(self.ptr), _ptr(a), a.nbytes, 0, e,
                                             1024) != 0:
            _raise(e)

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype, order="F")
        e = _err()
        if load_library().fastpls_hip_memcpy(_ptr(out), C.c_void_p(self.ptr), out.nbytes, 1, e,
                                             1024) != 0:
            _raise(e)
        return out

    def free(self):
        if self.ptr:
            e = _err()
            if load_library().fastpls_hip_device_free(C.c_void_p(self.ptr), e, 1024) != 0:
                _raise(e)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def latent_data_device(x_ptr, y_ptr, precision, n, p, q, factors=10, decay=1.0, noise=0.1,
                       seed=0, round32=False):
    """Fill caller-allocated device buffers with latent-factor data (round32:
    FP64 buffers hold the FP32-rounded values)."""
    e = _err()
    if load_library().fastpls_hip_latent_data(C.c_void_p(x_ptr), C.c_void_p(y_ptr), precision, n,
                                              p, q, factors, decay, noise, seed, int(round32), e,
                                              1024) != 0:
        _raise(e)


BATCH_FIELDS = {"R": 0, "Q": 1, "T": 2, "mX": 3, "vX": 4, "mY": 5}


def batch_splits(handle=None):
    """Split lengths of a batch, or the process default for new batches
    (fastpls_hip_batch_splits)."""
    out = (C.c_longlong * 4)()
    e = _err()
    if load_library().fastpls_hip_batch_splits(handle, out, e, 1024) != 0:
        _raise(e)
    return tuple(int(v) for v in out)


def set_batch_splits(lengths):
    """Replace the process default split lengths (fastpls_hip_batch_set_splits)."""
    arr = (C.c_longlong * 4)(*[int(v) for v in lengths])
    e = _err()
    if load_library().fastpls_hip_batch_set_splits(arr, e, 1024) != 0:
        _raise(e)


class SimplsBatch:
    """Many SIMPLS fits in one grouped execution (fastpls_hip_simpls_batch_create).

    ``groups``: list of (device pointer of X, n, p). ``fits``: list of dicts with
    keys group, q, components, seed and exactly one of y (device pointer of the
    n x q responses) / labels (device pointer of one-based labels), plus an
    optional perm (device pointer of a row permutation of the responses).
    """

    def __init__(self, precision, groups, fits, scaling=2, oversample=32, power=5, reorth=-1):
        self.precision = int(precision)
        self.dtype = np.float64 if self.precision == 64 else np.float32
        self.groups = [(int(g[1]), int(g[2])) for g in groups]
        self.fits = [dict(f) for f in fits]
        G, F = len(groups), len(fits)
        xs = (C.c_void_p * G)(*[C.c_void_p(g[0]) for g in groups])
        ns = (C.c_int * G)(*[g[1] for g in groups])
        ps = (C.c_int * G)(*[g[2] for g in groups])
        grp = (C.c_int * F)(*[f["group"] for f in fits])
        has_y = any(f.get("y") for f in fits)
        has_l = any(f.get("labels") for f in fits)
        has_p = any(f.get("perm") for f in fits)
        ys = (C.c_void_p * F)(*[C.c_void_p(f.get("y") or 0) for f in fits]) if has_y else None
        ls = (C.c_void_p * F)(*[C.c_void_p(f.get("labels") or 0) for f in fits]) if has_l else None
        pm = (C.c_void_p * F)(*[C.c_void_p(f.get("perm") or 0) for f in fits]) if has_p else None
        qs = (C.c_int * F)(*[f["q"] for f in fits])
        cs = (C.c_int * F)(*[f["components"] for f in fits])
        sd = (C.c_ulonglong * F)(*[int(f.get("seed", 1)) for f in fits])
        e = _err()
        self._h = load_library().fastpls_hip_simpls_batch_create(
            self.precision, G, C.cast(xs, C.c_void_p), ns, ps, F, grp,
            C.cast(ys, C.c_void_p) if ys is not None else None,
            C.cast(ls, C.c_void_p) if ls is not None else None,
            C.cast(pm, C.c_void_p) if pm is not None else None, qs, cs, sd, int(scaling),
            int(oversample), int(power), int(reorth), e, 1024)
        if not self._h:
            _raise(e)

    def rerun(self):
        """Execute every fit again on the current contents of the device inputs
        (fastpls_hip_batch_rerun: a captured graph of the whole batch)."""
        e = _err()
        if load_library().fastpls_hip_batch_rerun(self._h, e, 1024) != 0:
            _raise(e)

    def autotune(self, trials=2):
        """Choose the batch's split lengths by measurement
        (fastpls_hip_batch_autotune): {"splits": (step_rows, step_cols,
        wide_rows, wide_cols), "seconds", "start_seconds", "plans"}."""
        rep = np.zeros(7)
        e = _err()
        if load_library().fastpls_hip_batch_autotune(self._h, int(trials), _dptr(rep), 7, e,
                                                      1024) != 0:
            _raise(e)
        return {"splits": tuple(int(v) for v in rep[:4]), "seconds": float(rep[4]),
                "start_seconds": float(rep[5]), "plans": int(rep[6])}

    def splits(self):
        return batch_splits(self._h)

    def lanes(self):
        """Lane count of the batch (fastpls_hip_batch_lanes)."""
        out = C.c_int(0)
        e = _err()
        if load_library().fastpls_hip_batch_lanes(self._h, C.byref(out), e, 1024) != 0:
            _raise(e)
        return out.value

    def status(self):
        F = len(self.fits)
        done = (C.c_int * F)()
        flags = (C.c_int * F)()
        sec = C.c_double(0.0)
        e = _err()
        if load_library().fastpls_hip_batch_status(self._h, done, flags, C.byref(sec), e,
                                                   1024) != 0:
            _raise(e)
        return list(done), list(flags), sec.value

    PHASES = ("crossprod", "refresh", "score_valu", "score_mfma", "score_sums", "loading_valu",
              "loading_mfma", "update")

    def phases(self):
        """Device-time phases (set FASTPLS_BENCH_PHASE_TIMING=1 before creating
        the batch): {"seconds": s, phase: {"seconds", "bytes", "launches"}}."""
        out = np.zeros(1 + 3 * len(self.PHASES))
        e = _err()
        if load_library().fastpls_hip_batch_phases(self._h, _ptr(out), out.size, e, 1024) != 0:
            _raise(e)
        res = {"seconds": float(out[0])}
        for i, name in enumerate(self.PHASES):
            res[name] = {"seconds": float(out[1 + 3 * i]), "bytes": float(out[2 + 3 * i]),
                         "launches": int(out[3 + 3 * i])}
        return res

    def export(self, fit, field, f64=False):
        """Field of one fit in the batch precision, or (f64=True) the FP64
        values the fit holds."""
        f = BATCH_FIELDS[field] if isinstance(field, str) else int(field)
        n, p = self.groups[self.fits[fit]["group"]]
        q, a = self.fits[fit]["q"], self.fits[fit]["components"]
        shape = {0: (p, a), 1: (q, a), 2: (n, a), 3: (p,), 4: (p,), 5: (q,)}[f]
        out = np.empty(shape, dtype=np.float64 if f64 else self.dtype, order="F")
        e = _err()
        fn = (load_library().fastpls_hip_batch_export64 if f64
              else load_library().fastpls_hip_batch_export)
        if fn(self._h, int(fit), f, _ptr(out), out.size, e, 1024) != 0:
            _raise(e)
        return out

    def close(self):
        if getattr(self, "_h", None):
            load_library().fastpls_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _cv_regression(entry, x, y, folds, prefixes, precision, scaling, metric, oversample, power,
                   seed, store_predictions):
    dt = np.float64 if precision == 64 else np.float32
    xa = np.asfortranarray(x, dtype=dt)
    ya = np.asfortranarray(np.asarray(y).reshape(len(x), -1), dtype=dt)
    n, p = xa.shape
    q = ya.shape[1]
    fo = np.ascontiguousarray(folds, dtype=np.int32)
    pr = np.ascontiguousarray(prefixes, dtype=np.int32)
    k = len(pr)
    status = np.zeros(max(int(fo.max()), 1), dtype=np.int32)
    metrics, q2, rmsd, r2 = (np.zeros(k) for _ in range(4))
    preds = np.zeros((n, q, k), order="F") if store_predictions else None
    e = _err()
    rc = getattr(load_library(), entry)(
        _ptr(xa), _ptr(ya), _iptr(fo), int(precision), n, p, q, _iptr(pr), k, int(scaling),
        int(metric), int(oversample), int(power), int(seed), int(bool(store_predictions)),
        _dptr(preds), _iptr(status), _dptr(metrics), _dptr(q2), _dptr(rmsd), _dptr(r2), e, 1024)
    if rc != 0:
        _raise(e)
    return dict(status=status, metrics=metrics, q2=q2, rmsd=rmsd, observed_r2=r2,
                predictions=preds)


def simpls_cv_regression(x, y, folds, prefixes, precision=64, scaling=2, metric=3,
                         oversample=32, power=5, seed=1, store_predictions=True):
    """fastpls_hip_simpls_cv_regression: fold metrics for every prefix (all
    folds fitted in one batch, one shared seed)."""
    return _cv_regression("fastpls_hip_simpls_cv_regression", x, y, folds, prefixes, precision,
                          scaling, metric, oversample, power, seed, store_predictions)


def plssvd_cv_regression(x, y, folds, prefixes, precision=64, scaling=2, metric=3,
                         oversample=32, power=5, seed=1, store_predictions=True):
    """fastpls_hip_plssvd_cv_regression: fold metrics for every prefix (fold
    f fitted with seed + f)."""
    return _cv_regression("fastpls_hip_plssvd_cv_regression", x, y, folds, prefixes, precision,
                          scaling, metric, oversample, power, seed, store_predictions)


def _cv_classification(entry, x, labels, folds, classes, prefixes, precision, scaling, lda,
                       oversample, power, seed, store_predictions, store_scores):
    dt = np.float64 if precision == 64 else np.float32
    xa = np.asfortranarray(x, dtype=dt)
    n, p = xa.shape
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    fo = np.ascontiguousarray(folds, dtype=np.int32)
    pr = np.ascontiguousarray(prefixes, dtype=np.int32)
    k = len(pr)
    status = np.zeros(max(int(fo.max()), 1), dtype=np.int32)
    metrics = np.zeros(k)
    preds = np.zeros((n, k), dtype=np.int32, order="F") if store_predictions else None
    scores = np.zeros((n, int(classes), k), dtype=dt, order="F") if store_scores else None
    e = _err()
    rc = getattr(load_library(), entry)(
        _ptr(xa), _iptr(lab), _iptr(fo), int(precision), n, p, int(classes), _iptr(pr), k,
        int(scaling), int(bool(lda)), int(oversample), int(power), int(seed),
        int(bool(store_predictions)), int(bool(store_scores)), _iptr(preds), _ptr(scores),
        _iptr(status), _dptr(metrics), e, 1024)
    if rc != 0:
        _raise(e)
    return dict(status=status, metrics=metrics, predictions=preds, scores=scores)


def simpls_cv_classification(x, labels, folds, classes, prefixes, precision=64, scaling=2,
                             lda=False, oversample=32, power=5, seed=1, store_predictions=True,
                             store_scores=False):
    """fastpls_hip_simpls_cv_classification: accuracy per prefix."""
    return _cv_classification("fastpls_hip_simpls_cv_classification", x, labels, folds, classes,
                              prefixes, precision, scaling, lda, oversample, power, seed,
                              store_predictions, store_scores)


def plssvd_cv_classification(x, labels, folds, classes, prefixes, precision=64, scaling=2,
                             lda=False, oversample=32, power=5, seed=1, store_predictions=True,
                             store_scores=False):
    """fastpls_hip_plssvd_cv_classification: accuracy per prefix (fold f
    fitted with seed + f)."""
    return _cv_classification("fastpls_hip_plssvd_cv_classification", x, labels, folds, classes,
                              prefixes, precision, scaling, lda, oversample, power, seed,
                              store_predictions, store_scores)


SERVER_STATS = ("completed", "failed", "predict_p50_ms", "predict_p99_ms", "fit_p50_ms",
                "fit_p99_ms", "completed_per_s", "peak_bytes", "budget_bytes", "peak_concurrent",
                "predict_batches", "mean_batch_rows", "max_batch_requests", "span_s",
                "bytes_in_use", "resident_models", "fit_service_ms", "predict_service_ms")


class Server:
    """fastpls_hip_server_*: continuous batching of fit and predict requests
    under an HBM budget. Arrays handed to a request are kept alive here until
    the request is polled as finished."""

    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, budget_bytes: float, workers: int = 4, max_batch_rows: int = 8192,
                 max_concurrent_fits: int = 0):
        e = _err()
        h = load_library().fastpls_hip_server_create(float(budget_bytes), int(workers),
                                                     int(max_batch_rows),
                                                     int(max_concurrent_fits), e, 1024)
        if not h:
            _raise(e)
        self._h = h
        self._keep = {}
        self._meta = {}

    def submit_fit(self, x, y=None, labels=None, classes=None, method=SIMPLS, precision=64,
                   components=2, scaling=2, oversample=32, power=5, seed=1, device=None):
        """Host arrays, or device=(x_ptr, y_ptr or labels_ptr, n, p, q) for
        device-resident inputs. Returns the request id, which is also the id
        of the fitted model."""
        e = _err()
        lib = load_library()
        if device is not None:
            xp, rp, n, p, q = device
            yp, lp = (rp, None) if labels is None else (None, rp)
            keep = ()
            rid = lib.fastpls_hip_server_submit_fit(
                self._h, int(method), C.c_void_p(xp), C.c_void_p(yp) if yp else None,
                C.cast(C.c_void_p(lp), _ip) if lp else None, int(precision), 1, int(n), int(p),
                int(q), int(components), int(scaling), int(oversample), int(power), int(seed),
                e, 1024)
        else:
            dt = _dtype(precision)
            xa = _fmat(x, dt)
            n, p = xa.shape
            if labels is not None:
                la = np.ascontiguousarray(labels, dtype=np.int32)
                ya, q = None, int(classes)
            else:
                ya = _fmat(np.asarray(y).reshape(n, -1), dt)
                la, q = None, ya.shape[1]
            keep = (xa, ya, la)
            rid = lib.fastpls_hip_server_submit_fit(
                self._h, int(method), _ptr(xa), _ptr(ya), _iptr(la), int(precision), 0, n, p, q,
                int(components), int(scaling), int(oversample), int(power), int(seed), e, 1024)
        if rid == 0:
            _raise(e)
        self._keep[rid] = keep
        self._meta[rid] = (int(precision), int(q), int(components))
        return rid

    def submit_predict(self, model, x=None, prefix=None, device=None):
        """Host rows x (rows x p), or device=(ptr, ld, rows) for device rows.
        Returns (request id, output array filled when the request is done)."""
        precision, q, components = self._meta[model]
        prefix = components if prefix is None else int(prefix)
        dt = _dtype(precision)
        e = _err()
        lib = load_library()
        if device is not None:
            ptr, ld, rows = device
            out = np.zeros((int(rows), q), dtype=dt, order="F")
            rid = lib.fastpls_hip_server_submit_predict(self._h, int(model), C.c_void_p(ptr),
                                                        int(ld), 1, int(rows), prefix, _ptr(out),
                                                        e, 1024)
            keep = (out,)
        else:
            xa = _fmat(x, dt)
            rows = xa.shape[0]
            out = np.zeros((rows, q), dtype=dt, order="F")
            rid = lib.fastpls_hip_server_submit_predict(self._h, int(model), _ptr(xa), rows, 0,
                                                        rows, prefix, _ptr(out), e, 1024)
            keep = (xa, out)
        if rid == 0:
            _raise(e)
        self._keep[rid] = keep
        return rid, out

    def poll(self, request, wait=True):
        """(state, latency ms); raises HipError for a failed request."""
        state = C.c_int(0)
        lat = C.c_double(0.0)
        e = _err()
        if load_library().fastpls_hip_server_poll(self._h, int(request), int(bool(wait)),
                                                  C.byref(state), C.byref(lat), e, 1024) != 0:
            _raise(e)
        if state.value in (self.DONE, self.FAILED):
            self._keep.pop(request, None)
            if state.value == self.FAILED:
                _raise(e)
        return state.value, lat.value

    def release(self, model):
        e = _err()
        if load_library().fastpls_hip_server_release(self._h, int(model), e, 1024) != 0:
            _raise(e)

    def export(self, model, field, shape):
        precision = self._meta[model][0]
        out = np.zeros(shape, dtype=_dtype(precision), order="F")
        e = _err()
        f = FIELDS[field] if isinstance(field, str) else int(field)
        if load_library().fastpls_hip_server_export(self._h, int(model), f, _ptr(out), out.size,
                                                    e, 1024) != 0:
            _raise(e)
        return out

    def stats(self):
        out = np.zeros(len(SERVER_STATS))
        e = _err()
        if load_library().fastpls_hip_server_stats(self._h, _dptr(out), len(SERVER_STATS), e,
                                                   1024) != 0:
            _raise(e)
        return dict(zip(SERVER_STATS, out.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            load_library().fastpls_hip_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- fastsvd / fastcor
AUDIT_FIELDS = ("certified", "deterministic_fallback", "attempts", "effective_oversample",
                "effective_power", "effective_seed", "subspace_error", "singular_value_error",
                "triplet_residual", "omitted_direction_ratio")


def fastsvd(x, ncomp, oversample=32, power=5, seed=1, left_only=False, precision=64):
    """fastsvd(x, ncomp, oversample, power, seed) (R/main.R:8181-8204): the
    audited operator rSVD of x on the device (audited_rsvd.hpp:298-414).
    Returns dict(U (m x k), s (k), Vt (k x n, None when left_only), audit)."""
    lib = load_library()
    a = _fmat(x, _dtype(precision))
    m, n = a.shape
    k = min(int(ncomp), m, n)
    U = np.zeros((m, max(k, 1)), order="F")
    s = np.zeros(max(k, 1))
    Vt = None if left_only else np.zeros((max(k, 1), n), order="F")
    got = C.c_int(0)
    audit = np.zeros(10)
    e = _err()
    if lib.fastpls_hip_rsvd(_ptr(a), precision, 0, m, n, int(ncomp), int(oversample), int(power),
                            C.c_uint(int(seed) & 0xFFFFFFFF), int(bool(left_only)), _dptr(U),
                            _dptr(s), _dptr(Vt), C.byref(got), _dptr(audit), e, 1024) != 0:
        _raise(e)
    g = got.value
    return dict(U=U[:, :g], s=s[:g], Vt=None if left_only else np.asfortranarray(Vt[:g]),
                audit=dict(zip(AUDIT_FIELDS, audit.tolist())))


def fastcor(a, b=None, by_row=True, diagonal=True, precision=64):
    """fastcor(a, b, byrow = TRUE, diag = TRUE) (R/main.R:12799,
    src/r_api.cpp:4054-4166): Pearson correlations of the units (rows when
    by_row, else columns) of a with those of b, or only the matching pairs
    when diagonal; without b, the full correlation of a's units."""
    lib = load_library()
    dt = _dtype(precision)
    aa = _fmat(a, dt)
    bb = None if b is None else _fmat(b, dt)
    ua = aa.shape[0] if by_row else aa.shape[1]
    if bb is None:
        ub, diagonal = ua, False
    else:
        ub = bb.shape[0] if by_row else bb.shape[1]
    out = np.zeros(ua) if diagonal else np.zeros((ua, ub), order="F")
    e = _err()
    if lib.fastpls_hip_fastcor(_ptr(aa), aa.shape[0], aa.shape[1], _ptr(bb),
                               0 if bb is None else bb.shape[0], 0 if bb is None else bb.shape[1],
                               precision, 0, int(bool(by_row)), int(bool(diagonal)), _dptr(out),
                               e, 1024) != 0:
        _raise(e)
    return out


def _view32(a):
    """(pointer, rows, columns, leading dimension) of a 2-D float32 array that
    is column-major with unit row stride (a column slice of a larger Fortran
    array keeps its parent's leading dimension)."""
    if a.dtype != np.float32 or a.ndim != 2:
        raise ValueError("expected a column-major float32 matrix view")
    if a.size == 0:
        return _ptr(a), a.shape[0], a.shape[1], max(a.shape[0], 1)
    if a.strides[0] != 4 and a.shape[0] > 1:
        raise ValueError("expected a column-major float32 matrix view")
    ld = a.strides[1] // 4 if a.shape[1] > 1 else max(a.shape[0], 1)
    return _ptr(a), a.shape[0], a.shape[1], ld


class CoreWorkspace:
    """The FP32 core hooks (fastpls_hip_core_*_f32) over host matrix views."""

    def __init__(self):
        e = _err()
        self._h = load_library().fastpls_hip_core_workspace_create_f32(e, 1024)
        if not self._h:
            _raise(e)

    def gemm_into(self, left, right, transpose_left, transpose_right, out):
        lp, lr, lc, ll = _view32(left)
        rp, rr, rc, rl = _view32(right)
        op, orows, ocols, old = _view32(out)
        e = _err()
        if load_library().fastpls_hip_core_gemm_into_f32(
                self._h, lp, lr, lc, ll, rp, rr, rc, rl, int(bool(transpose_left)),
                int(bool(transpose_right)), op, orows, ocols, old, e, 1024) != 0:
            _raise(e)

    def self_gram_into(self, a, transpose_input, out, full_output):
        ap, ar, ac, al = _view32(a)
        op, orows, ocols, old = _view32(out)
        e = _err()
        if load_library().fastpls_hip_core_self_gram_into_f32(
                self._h, ap, ar, ac, al, int(bool(transpose_input)), op, orows, ocols, old,
                int(bool(full_output)), e, 1024) != 0:
            _raise(e)

    def close(self):
        if self._h:
            load_library().fastpls_hip_core_workspace_destroy_f32(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def core_gemm_f32(left, right, transpose_left=False, transpose_right=False):
    """cuda_core_gemm_f32 counterpart: op(left) op(right) as a new float32 array."""
    lp, lr, lc, ll = _view32(left)
    rp, rr, rc, rl = _view32(right)
    rows = lc if transpose_left else lr
    cols = rr if transpose_right else rc
    out = np.zeros((rows, cols), dtype=np.float32, order="F")
    e = _err()
    if load_library().fastpls_hip_core_gemm_f32(lp, lr, lc, ll, rp, rr, rc, rl,
                                                int(bool(transpose_left)),
                                                int(bool(transpose_right)), _ptr(out), e,
                                                1024) != 0:
        _raise(e)
    return out


def lda_moments(gram, class_sums, counts, n, prefix):
    """LDA from score moments on the device (fastpls_hip_lda_moments):
    returns dict(linear=K x prefix, constants=K, ridge)."""
    g = np.asfortranarray(gram, dtype=np.float64)
    cs = np.asfortranarray(class_sums, dtype=np.float64)
    cnt = np.ascontiguousarray(counts, dtype=np.int32)
    a, K = g.shape[0], cs.shape[0]
    lin = np.zeros((K, prefix), order="F"); con = np.zeros(K); ridge = np.zeros(1)
    e = _err()
    if load_library().fastpls_hip_lda_moments(_dptr(g), a, _dptr(cs), _iptr(cnt), K, int(n),
                                              int(prefix), _dptr(lin), _dptr(con), _dptr(ridge),
                                              e, 1024) != 0:
        _raise(e)
    return dict(linear=lin, constants=con, ridge=float(ridge[0]))


def plssvd_weights(gram, sigma, Q):
    """PLS-SVD prefix weights on the device (fastpls_hip_plssvd_weights):
    list of W_k (k x q) for k = 1..components."""
    g = np.asfortranarray(gram, dtype=np.float64)
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    q_ = np.asfortranarray(Q, dtype=np.float64)
    A, q = g.shape[0], q_.shape[0]
    W = np.zeros(A * A * q)
    e = _err()
    if load_library().fastpls_hip_plssvd_weights(_dptr(g), _dptr(s), _dptr(q_), A, q, _dptr(W), e,
                                                 1024) != 0:
        _raise(e)
    return [W[(k - 1) * A * q:(k - 1) * A * q + k * q].reshape((k, q), order="F")
            for k in range(1, A + 1)]
