"""The no-allocation prediction entry (GPU, through the C ABI).

fastpls_hip_predict_noalloc takes device rows and a device output and keeps
every intermediate in the library's per-GPU scratch array. It must give what
fastpls_hip_predict_path gives for the same model, rows and prefix (relative
1e-12 at precision 64, 1e-6 at precision 32, i.e. FP32 output rounding), on
two calls in a row, the second with fewer rows and another prefix (the
scratch keeps the first call's contents), for SIMPLS and PLS-SVD models (the
PLS-SVD FP32 route also keeps its bf16x6 operand workspace in the scratch).
A request larger than the scratch fails with an error before any launch.
"""
import numpy as np
import pytest

from conftest import rel
from oracle.oracle import latent_data

pytestmark = pytest.mark.gpu

TOL = {64: 1e-12, 32: 1e-6}


@pytest.mark.parametrize("method,precision", [(0, 32), (0, 64), (1, 32), (1, 64)])
def test_noalloc_predict_matches_predict_path(hip, method, precision):
    fp = hip[0]
    dt = np.float64 if precision == 64 else np.float32
    X, Y = latent_data(3000, 700, 4, factors=8, noise=0.1, seed=41)
    A = 8 if method == 0 else 4
    m = fp.ResidentModel.fit(X.astype(dt), Y.astype(dt), method=method, precision=precision,
                             components=A, scaling=2, oversample=10, power=2, seed=3)
    Xq, _ = latent_data(2500, 700, 4, factors=8, noise=0.1, seed=42)
    Xq = np.asfortranarray(Xq.astype(dt))
    for rows, prefix in [(2500, A), (600, 3)]:
        print(f"method {method} precision {precision}: rows {rows} prefix {prefix}", flush=True)
        xr = np.asfortranarray(Xq[:rows])
        xb = fp.DeviceBuffer.from_array(xr)
        ob = fp.DeviceBuffer(rows * 4 * np.dtype(dt).itemsize)
        m.predict_noalloc(xb.ptr, rows, prefix, ob.ptr)
        got = ob.download((rows, 4), dt)
        want = m.predict(xr, prefix)
        assert rel(got, want) <= TOL[precision], (rows, prefix, rel(got, want))
        xb.free(); ob.free()
    m.close()


def test_noalloc_rejects_requests_beyond_the_scratch(hip):
    fp = hip[0]
    X, Y = latent_data(400, 50, 2, factors=4, seed=43)
    m = fp.ResidentModel.fit(X, Y, precision=64, components=4, scaling=2, oversample=4, power=1)
    xb = fp.DeviceBuffer.from_array(np.asfortranarray(X[:10]))
    ob = fp.DeviceBuffer(10 * 2 * 8)
    # 4e6 rows x 4 scores alone exceed the 4 Mi-double scratch: the call fails
    # on the host, before any launch reads the (10-row) buffers
    with pytest.raises(fp.HipError, match="scratch doubles"):
        m.predict_noalloc(xb.ptr, 4_000_000, 4, ob.ptr)
    xb.free(); ob.free(); m.close()
