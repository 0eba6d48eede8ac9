// Internal GEMM launchers shared by the density and contraction drivers.
#pragma once
#include "xr_common.h"

// FP64 MFMA GEMM (no locking; callers hold the context lock).
// ta = 1: A is K x M (A[k*lda+m]); tb = 1: B is N x K (B[n*ldb+k]).
int xr_dgemm_launch(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                    const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb,
                    int64_t sB, double beta, double* C, int64_t ldc, int64_t sC, int64_t batch);

// C = alpha A^T A (A: K x n, C: n x n): upper tiles, then the mirror.
int xr_dgemm_syrk_tn(xr_ctx* c, int64_t n, int64_t K, double alpha, const double* A, int64_t lda, double* C,
                     int64_t ldc);
int xr_dgemm_syrk_nt(xr_ctx* c, int64_t n, int64_t K, double alpha, const double* A, int64_t lda, double* C,
                     int64_t ldc);

// Split-FP32 product (hi/lo FP32 MFMA, FP64 accumulation) and the
// precision-switched GEMM used by the two-electron contractions.
int xr_dgemm_split_launch(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                          const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb,
                          int64_t sB, double beta, double* C, int64_t ldc, int64_t sC, int64_t batch);
// Split-int8 product (Ozaki digit planes on the int8 MFMA).
int xr_dgemm_i8_launch(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
                       const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB,
                       double beta, double* C, int64_t ldc, int64_t sC, int64_t batch);
bool xr_i8_pays(int64_t M, int64_t N, int64_t K, int64_t batch);
int xr_gemm_2e(xr_ctx* c, int ta, int tb, int64_t M, int64_t N, int64_t K, double alpha,
               const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB,
               double beta, double* C, int64_t ldc, int64_t sC, int64_t batch);
