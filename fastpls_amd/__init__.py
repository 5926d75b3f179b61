"""fastpls_amd: MI355X (gfx950) HIP backend for fastPLS.

The compute path is the C ABI library ``lib/libfastpls_hip.so`` (declared in
``include/fastpls_hip.h``); this package is its Python binding.
"""
from .resident import (AUDIT_FIELDS, FIELDS, KERNELPLS, OPLS, PLSSVD, SIMPLS, CoreWorkspace, DeviceBuffer, HipError,
                       ResidentModel, Server, SimplsBatch, batch_splits, core_gemm_f32,
                       device_info, fastcor, set_batch_splits,
                       fastsvd, gemm, lda_moments, plssvd_weights,
                       latent_data_device, load_library,
                       plssvd_cv_classification, plssvd_cv_regression, simpls_cv_classification,
                       simpls_cv_regression)

__all__ = ["AUDIT_FIELDS", "FIELDS", "KERNELPLS", "OPLS", "PLSSVD", "SIMPLS", "CoreWorkspace", "DeviceBuffer", "HipError",
           "ResidentModel", "Server", "SimplsBatch", "batch_splits", "core_gemm_f32",
           "device_info", "fastcor", "set_batch_splits",
           "fastsvd", "gemm", "lda_moments", "plssvd_weights",
           "latent_data_device", "load_library",
           "plssvd_cv_classification", "plssvd_cv_regression", "simpls_cv_classification",
           "simpls_cv_regression"]
