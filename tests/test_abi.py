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


# The reference's resident ABI (src/cuda_resident_api.h:11-77), restated as
# the interface a caller such as src/cuda_resident_r.cpp is written against:
# the same names, return types and parameter lists.
REFERENCE_PROTOTYPES = r"""
int fastpls_cuda_gemm(const void*, const void*, int, int, int, int, void*, char*, size_t);
int fastpls_resident_simpls_cv_classification(const void*, const int*, const int*, int, int,
    int, int, const int*, int, int, int, int, int, unsigned long long, int, int, int*, void*,
    int*, double*, char*, size_t);
int fastpls_resident_simpls_cv_regression(const void*, const void*, const int*, int, int, int,
    int, const int*, int, int, int, int, int, unsigned long long, int, double*, int*, double*,
    double*, double*, double*, char*, size_t);
int fastpls_resident_plssvd_cv_classification(const void*, const int*, const int*, int, int,
    int, int, const int*, int, int, int, int, int, unsigned long long, int, int, int*, void*,
    int*, double*, char*, size_t);
int fastpls_resident_plssvd_cv_regression(const void*, const void*, const int*, int, int, int,
    int, const int*, int, int, int, int, int, unsigned long long, int, double*, int*, double*,
    double*, double*, double*, char*, size_t);
void* fastpls_resident_simpls_create(const void*, const void*, const int*, int, int, int, int,
    int, int, int, int, int, unsigned long long, char*, size_t);
void* fastpls_resident_plssvd_create(const void*, const void*, const int*, int, int, int, int,
    int, int, int, int, int, unsigned long long, char*, size_t);
void* fastpls_resident_opls_create(const void*, const void*, const int*, int, int, int, int,
    int, int, int, int, int, unsigned long long, int, char*, size_t);
void* fastpls_resident_kernelpls_create(const void*, const void*, const int*, int, int, int,
    int, int, int, int, int, int, unsigned long long, int, double, int, double, char*, size_t);
int fastpls_resident_predict_path(void*, const void*, int, const int*, int, int, void*, char*,
    size_t);
void fastpls_resident_simpls_destroy(void*);
int fastpls_resident_export(void*, int, void*, size_t, char*, size_t);
int fastpls_resident_classify_path(void*, const void*, int, const int*, int, int, int, int*,
    char*, size_t);
int fastpls_resident_classify_response_path(void*, const void*, int, const int*, int, int, int,
    int*, void*, char*, size_t);
int fastpls_resident_response_sums(void*, const void*, const void*, const int*, int, int, void*,
    char*, size_t);
int fastpls_resident_project(void*, const void*, int, int, void*, char*, size_t);
int fastpls_resident_controls(void*, int*, int*, int*, int*, int*, int*, char*, size_t);
int fastpls_resident_compact(void*, int, char*, size_t);
"""


def test_reference_named_caller_links_through_the_alias_header(tmp_path):
    """A C++ caller written against the reference's names compiles with the
    alias header (every reference prototype, redeclared after it, must match
    the HIP declaration it maps to, or the compiler rejects the conflicting
    extern "C" declaration) and links against libfastpls_hip.so."""
    if not os.path.exists(LIB):
        import fastpls_amd.build as b
        b.build()
    names = re.findall(r"\b(fastpls_(?:cuda|resident)_[a-z_]+)\s*\(", REFERENCE_PROTOTYPES)
    assert len(names) == 18
    src = tmp_path / "caller.cpp"
    src.write_text(
        "#include <cstddef>\n#include \"fastpls_resident_alias.h\"\n"
        "extern \"C\" {\n" + REFERENCE_PROTOTYPES + "}\n"
        "int main() {\n  void* table[] = {" + ", ".join(f"(void*)&{n}" for n in names) + "};\n"
        "  return table[0] == nullptr;\n}\n")
    exe = tmp_path / "caller"
    subprocess.run(["g++", "-std=c++17", "-Werror", f"-I{os.path.join(ROOT, 'include')}", str(src),
                    f"-L{os.path.dirname(LIB)}", "-lfastpls_hip",
                    f"-Wl,-rpath,{os.path.dirname(LIB)}", "-Wl,--unresolved-symbols=ignore-in-shared-libs",
                    "-o", str(exe)], check=True)
