// MFMA GEMM C[m][n] = alpha * sum_k A[m][k] * B[n][k] (A complex, B complex or real, C complex),
// the shape of both Cholesky contractions of a walker step:
//   vbias  (Real3IndexFactorization.hpp:350-361): A = G [W][N*M], B = Lnak [K][N*M]  (complex)
//   vHS    (Real3IndexFactorization.hpp:297-313): A = X [W][K],   B = Likn [M*M][K]   (real)
//
// Tiling for CDNA4: a 256-thread workgroup owns a 64x64 tile of C, each 64-wide wavefront a
// 32x32 sub-tile made of 2x2 MFMA 16x16x4 tiles with separate real/imag accumulators
// (4 real MFMAs per complex tile and k-step, 2 when B is real). Along k, lane group g = lane/16
// takes k = k0 + 4g + s in MFMA step s, so each lane reads 4 consecutive complex numbers
// (64 bytes) per 16-wide k chunk; A and B use the same permutation, which leaves the sum
// unchanged. Both FP64 (16x16x4f64) and FP32 (16x16x4f32) MFMA are instantiated; the FP32 form
// converts A on load and reads B from the FP32 copy of the Cholesky tensor.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kTile = 64;

template <typename R> struct cplx2 { R re, im; };

template <typename R, bool BC>
__global__ __launch_bounds__(256) void k_gemm_nt(int m, int n, int k, double alpha,
                                                 const double* __restrict__ A, long lda,
                                                 const R* __restrict__ B, long ldb,
                                                 double* __restrict__ C, long ldc) {
  typedef typename acc_of<R>::type acc_t;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int m0 = blockIdx.x * kTile + (wave >> 1) * 32;
  const int n0 = blockIdx.y * kTile + (wave & 1) * 32;

  acc_t cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = acc_t{0, 0, 0, 0};
      ci[a][b] = acc_t{0, 0, 0, 0};
    }

  for (int k0 = 0; k0 < k; k0 += 16) {
    const int kb = k0 + 4 * g;
    R ar[2][4], ai[2][4], br[2][4], bi[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int row = m0 + t * 16 + r16;
      const double* pa = A + 2 * ((long)row * lda + kb);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bool ok = row < m && kb + s < k;
        ar[t][s] = ok ? (R)pa[2 * s] : (R)0;
        ai[t][s] = ok ? (R)pa[2 * s + 1] : (R)0;
      }
      const int col = n0 + t * 16 + r16;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bool ok = col < n && kb + s < k;
        if (BC) {
          const R* pb = B + 2 * ((long)col * ldb + kb);
          br[t][s] = ok ? pb[2 * s] : (R)0;
          bi[t][s] = ok ? pb[2 * s + 1] : (R)0;
        } else {
          const R* pb = B + ((long)col * ldb + kb);
          br[t][s] = ok ? pb[s] : (R)0;
          bi[t][s] = (R)0;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          cr[a][b] = mfma16x16x4(ar[a][s], br[b][s], cr[a][b]);
          ci[a][b] = mfma16x16x4(ai[a][s], br[b][s], ci[a][b]);
          if (BC) {
            cr[a][b] = mfma16x16x4(-ai[a][s], bi[b][s], cr[a][b]);
            ci[a][b] = mfma16x16x4(ar[a][s], bi[b][s], ci[a][b]);
          }
        }
  }

#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = n0 + b * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + a * 16 + acc_row<R>(g, r);
        if (row < m && col < n) {
          double* pc = C + 2 * ((long)row * ldc + col);
          pc[0] = alpha * (double)cr[a][b][r];
          pc[1] = alpha * (double)ci[a][b][r];
        }
      }
    }
}

}  // namespace

void launch_gemm_nt_z(hipStream_t s, int precision, int m, int n, int k, double alpha,
                      const double* A, long lda, const void* B, long ldb, bool b_complex,
                      double* C, long ldc) {
  if (m <= 0 || n <= 0) return;
  dim3 grid((m + kTile - 1) / kTile, (n + kTile - 1) / kTile);
  if (precision == AFQ_FP64) {
    if (b_complex)
      hipLaunchKernelGGL((k_gemm_nt<double, true>), grid, dim3(256), 0, s, m, n, k, alpha, A,
                         lda, (const double*)B, ldb, C, ldc);
    else
      hipLaunchKernelGGL((k_gemm_nt<double, false>), grid, dim3(256), 0, s, m, n, k, alpha, A,
                         lda, (const double*)B, ldb, C, ldc);
  } else {
    if (b_complex)
      hipLaunchKernelGGL((k_gemm_nt<float, true>), grid, dim3(256), 0, s, m, n, k, alpha, A,
                         lda, (const float*)B, ldb, C, ldc);
    else
      hipLaunchKernelGGL((k_gemm_nt<float, false>), grid, dim3(256), 0, s, m, n, k, alpha, A,
                         lda, (const float*)B, ldb, C, ldc);
  }
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
