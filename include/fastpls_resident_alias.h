/* Reference resident C ABI names on the HIP library.
 *
 * The reference's resident engine exports its C ABI under the names of
 * src/cuda_resident_api.h:11-77 (fastpls_cuda_gemm, fastpls_resident_*);
 * libfastpls_hip.so exports the same entry points, with identical parameter
 * lists and conventions, as fastpls_hip_* (fastpls_hip.h). Including this
 * header after, or in place of, the reference's declarations maps every
 * reference name onto its HIP entry, so an unchanged caller written against
 * the reference ABI (the R glue src/cuda_resident_r.cpp, built with
 * -DFASTPLS_HAS_CUDA -include fastpls_resident_alias.h) links against
 * libfastpls_hip.so. */
#ifndef FASTPLS_RESIDENT_ALIAS_H
#define FASTPLS_RESIDENT_ALIAS_H

#include "fastpls_hip.h"

#define fastpls_cuda_gemm fastpls_hip_gemm
#define fastpls_resident_simpls_cv_classification fastpls_hip_simpls_cv_classification
#define fastpls_resident_simpls_cv_regression fastpls_hip_simpls_cv_regression
#define fastpls_resident_plssvd_cv_classification fastpls_hip_plssvd_cv_classification
#define fastpls_resident_plssvd_cv_regression fastpls_hip_plssvd_cv_regression
#define fastpls_resident_simpls_create fastpls_hip_simpls_create
#define fastpls_resident_plssvd_create fastpls_hip_plssvd_create
#define fastpls_resident_opls_create fastpls_hip_opls_create
#define fastpls_resident_kernelpls_create fastpls_hip_kernelpls_create
#define fastpls_resident_predict_path fastpls_hip_predict_path
#define fastpls_resident_simpls_destroy fastpls_hip_destroy
#define fastpls_resident_export fastpls_hip_export
#define fastpls_resident_classify_path fastpls_hip_classify_path
#define fastpls_resident_classify_response_path fastpls_hip_classify_response_path
#define fastpls_resident_response_sums fastpls_hip_response_sums
#define fastpls_resident_project fastpls_hip_project
#define fastpls_resident_controls fastpls_hip_controls
#define fastpls_resident_compact fastpls_hip_compact

#endif
