// Declarations shared by the mixed-precision sources (mp_gemm.hip, mixed.hip).
#pragma once

#include "zmlhip_internal.hpp"
#include "../../include/zmlhip.h"

namespace zmlhip {

// C = beta C + alpha rs_i cs_j sum_k A_ik B_kj, FP32 accumulate (mp_gemm.hip).
void mp_gemm_f16(int64_t m, int64_t n, int64_t K, float alpha, const _Float16* A, int64_t lda,
                 const _Float16* BT, int64_t ldb, float beta, const float* rs, const float* cs,
                 float* C, int64_t ldc, hipStream_t s, int lower = 0);
void mp_gemm_bf16(int64_t m, int64_t n, int64_t K, float alpha, const __bf16* A, int64_t lda,
                  const __bf16* BT, int64_t ldb, float beta, const float* rs, const float* cs,
                  float* C, int64_t ldc, hipStream_t s, int lower = 0);
void mp_gemm_f32(int64_t m, int64_t n, int64_t K, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, const float* rs, const float* cs, float* C,
                 int64_t ldc, hipStream_t s, int lower = 0);
// C = alpha A B + beta C (row-major FP32) for few output tiles and a long K:
// K is split over up to 8 slices whose partial products (part: the floats
// mp_gemm_f32_splitk_floats gives) are summed in order by a second kernel
int64_t mp_gemm_f32_splitk_floats(int64_t m, int64_t n, int64_t K);
void mp_gemm_f32_splitk(int64_t m, int64_t n, int64_t K, float alpha, const float* A, int64_t lda, const float* B,
                        int64_t ldb, float beta, float* C, int64_t ldc, float* part, hipStream_t s);

// Precision of the trailing update of the mixed factorizations.
enum LowPrec : int { kPrecF32 = 0, kPrecF16 = 1, kPrecBF16 = 2 };

// Row-major LU machinery shared with the FP64 driver (getrf_f64.hip).
template <typename T>
struct LuOps {
  // trailing update C -= L U (rows x cols, inner kb)
  void (*gemm)(int64_t rows, int64_t cols, int64_t kb, const T* L, int64_t ldl, const T* U,
               int64_t ldu, T* C, int64_t ldc, hipStream_t s);
  // B (kb x cols) <- L11^-1 B, L11 unit lower (kb <= 64), inside a block column
  void (*trsm)(int64_t kb, int64_t cols, const T* L11, int64_t ld, T* B, int64_t ldb,
               hipStream_t s);
  // inv (kb x kb, ld kb) <- L11^-1 for the block column's unit lower L11
  void (*invert)(int64_t kb, const T* L11, int64_t ld, T* inv, hipStream_t s);
  // B (kb x cols) <- inv * B, through tmp (kb x cols, ld cols)
  void (*apply_inv)(int64_t kb, int64_t cols, const T* inv, T* B, int64_t ldb, T* tmp,
                    hipStream_t s);
  // update inside a block column (C -= L U); the trailing update when null
  void (*panel_gemm)(int64_t rows, int64_t cols, int64_t kb, const T* L, int64_t ldl, const T* U,
                     int64_t ldu, T* C, int64_t ldc, hipStream_t s);
  // optional: factor block column [j0, j0 + jb) (rows j0..m-1) with its pivots
  // taken inside the diagonal block; returns false, with the block column
  // unchanged, when that would not be the partial-pivoting result
  bool (*try_panel)(int64_t m, T* A, int64_t lda, int64_t j0, int64_t jb, int32_t* ipiv,
                    hipStream_t s);
  // optional: where L11^-1 of the block column starting at j0 is kept (the
  // driver's two alternating buffers when null)
  T* (*inv_store)(int64_t j0);
  // > 0: run the trailing updates on a stream that leaves this many CUs free,
  // so that the panel stream's workgroups always find empty CUs
  int spare_cus;
  // optional: the composed interchanges of block column j0 when the panel
  // already formed them (the driver composes them itself when null)
  const int32_t* (*block_perm)(int64_t j0);
  // optional, with try_panel: then try_panel only queues its attempt (its
  // return value says whether it made one), and this reports the outcome,
  // waiting for the panel stream; a failed attempt leaves the block column as
  // it was. The driver queues the previous step's trailing update in between.
  bool (*try_panel_finish)(int64_t j0, int64_t jb, T* A, int64_t lda, hipStream_t s) = nullptr;
  // optional: factors block column [j0, j0 + jb) in one launch, leaves it in
  // LAPACK row order and provides its composed interchanges through
  // block_perm (the leaf-by-leaf factor_panel when null). It is used for the
  // block columns with more than panel_min_rows rows (m - j0); the others
  // take factor_panel, and their interchanges are composed by the driver.
  void (*panel)(int64_t m, T* A, int64_t lda, int64_t j0, int64_t jb, int32_t* ipiv, int* info,
                hipStream_t s) = nullptr;
  int64_t panel_min_rows = 0;
};

int dposv_mixed(int layout, int uplo, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                const double* B, int64_t ldb, double* X, int64_t ldx, const zmlhip_ir_opts* o,
                zmlhip_ir_result* res, hipStream_t s);

int dgesv_mixed(int layout, int64_t n, int64_t nrhs, const double* A, int64_t lda, const double* B,
                int64_t ldb, double* X, int64_t ldx, const zmlhip_ir_opts* o, zmlhip_ir_result* res,
                hipStream_t s);

// Row interchanges of a block of rows (getrf_f64.hip): compose ipiv[k1..k2]
// into moves, apply the moves to `ncols` columns starting at `a`. A move list
// (count, then destination/source row pairs) lives in a workspace slot's
// buffer, perm_pairs_buffer(slot), of up to 2 x 512 moves.
int32_t* perm_pairs_buffer(int slot);
int32_t* compose_perm(int k1, int k2, const int32_t* ipiv, hipStream_t s, int slot);
template <typename T>
void apply_perm(int64_t ncols, T* a, int64_t lda, const int32_t* pairs, hipStream_t s);

template <typename T>
void getrf_lookahead(int64_t m, int64_t n, T* A, int64_t lda, int32_t* ipiv, int* info, int64_t nb,
                     const LuOps<T>& ops, hipStream_t s);

}  // namespace zmlhip
