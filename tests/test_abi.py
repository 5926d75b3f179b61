"""The C ABI library loads and exports every symbol declared in include/*.h
(CPU only: no compute call is made here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastpls_hip.h")
LIB = os.path.join(ROOT, "fastpls_amd", "lib", "libfastpls_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fastpls_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_surface():
    names = declared()
    # counterparts of src/cuda_resident_api.h:11-77 (CV drivers are batched elsewhere)
    for ref in ["gemm", "simpls_create", "plssvd_create", "opls_create", "kernelpls_create",
                "predict_path", "destroy", "export", "classify_path",
                "classify_response_path", "response_sums", "project", "controls", "compact"]:
        assert f"fastpls_hip_{ref}" in names, ref


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        import fastpls_amd.build as b
        b.build()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (fastpls_hip_[a-z0-9_]+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_binding_signatures_cover_the_header():
    from fastpls_amd.resident import SIGNATURES
    assert sorted(SIGNATURES) == declared()


def test_library_loads_and_binds():
    import fastpls_amd
    lib = fastpls_amd.load_library()
    for name in declared():
        assert getattr(lib, name) is not None


def test_no_cpu_fallback_without_a_device():
    import fastpls_amd
    try:
        fastpls_amd.device_info()
        has_gpu = True
    except fastpls_amd.HipError:
        has_gpu = False
    if has_gpu:
        pytest.skip("a device is present; the no-device error path is exercised on CPU hosts")
    with pytest.raises(fastpls_amd.HipError, match="no HIP device"):
        fastpls_amd.ResidentModel.fit([[1.0, 2.0], [3.0, 5.0], [1.0, 0.0]], [[1.0], [2.0], [0.0]],
                                      components=1)
