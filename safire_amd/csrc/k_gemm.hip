// MFMA GEMM C[m][n] = alpha * sum_k A[m][k] * B[n][k] (A complex, B complex or real, C complex),
// the shape of both Cholesky contractions of a walker step:
//   vbias  (Real3IndexFactorization.hpp:350-361): A = G [W][N*M], B = Lnak [K][N*M]  (complex)
//   vHS    (Real3IndexFactorization.hpp:297-313): A = X [W][K],   B = Likn [M*M][K]   (real)
//
// Tiling for CDNA4: a 256-thread workgroup owns a 64 x TN tile of C (TN = 64 in every launch; the
// template also takes 128), each 64-wide wavefront a 32 x TN/2 sub-tile made of 2 x TN/32 MFMA
// 16x16x4 tiles with separate real/imag accumulators (4 real MFMAs per complex tile and k-step,
// 2 when B is real). The workgroup stages each 64 x KS slice of A and TN x KS slice of B in LDS
// once (every thread loads 1/256 of it; rows padded to KS + 1
// elements so that the 16 rows of an MFMA fragment fall into different banks), double-buffered:
// the next slice's loads are in flight in registers while the current slice is multiplied, with
// one barrier per slice. Along k, lane group g = lane/16 takes k = k0 + 4g + s in MFMA step s;
// A and B use the same permutation, which leaves the sum unchanged. Both FP64 (16x16x4f64) and
// FP32 (16x16x4f32) MFMA are instantiated; the FP32 form converts A when it is staged and reads
// B from the FP32 copy of the Cholesky tensor.
//
// vbias has few output tiles (W x K = 512 x 320 is 40 tiles) and a long k (N*M), so k is split
// over grid.z: each split writes its partial tile to a workspace and a second kernel sums the
// splits in a fixed order (deterministic, no atomics) and applies alpha.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kTile = 64;

template <typename R, bool BC, int KS, int TN>
__global__ __launch_bounds__(256) void k_gemm_nt(int m, int n, int k, int kchunk, double alpha,
                                                 const double* __restrict__ A, long lda,
                                                 const R* __restrict__ B, long ldb,
                                                 double* __restrict__ C, long ldc,
                                                 long split_stride, int nmajor) {
  typedef typename acc_of<R>::type acc_t;
  constexpr int LDK = KS + 1;          // padded k stride of a staged row (elements)
  constexpr int NB = BC ? 2 : 1;       // R values per element of B
  constexpr int KPT = KS / 4;          // k values each thread stages per row and slice
  constexpr int NTW = TN / 32;         // 16-wide MFMA column tiles per wavefront
  constexpr int BROWS = TN / 64;       // B rows each thread stages per slice
  __shared__ R sA[2][kTile * LDK * 2];
  __shared__ R sB[2][TN * LDK * NB];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * (TN / 2);
  // nmajor: the column tile is the fastest grid index, so consecutive workgroups (which the
  // dispatcher deals round-robin to the 8 XCDs) share A tiles and each XCD's L2 holds 1/8 of B
  const int tile_m = nmajor ? blockIdx.y : blockIdx.x;
  const int tile_n = nmajor ? blockIdx.x : blockIdx.y;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(k, kbeg + kchunk);
  C += blockIdx.z * split_stride;

  // staging share of this thread: row srow of the A and B tiles, k = k0 + skq .. + KPT - 1
  const int srow = tid >> 2, skq = (tid & 3) * KPT;
  const int gm = tile_m * kTile + srow;
  R pa[2 * KPT], pb[BROWS][NB * KPT];
  auto gload = [&](int k0) {
    const int kb = k0 + skq;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const bool ok = kb + j < kend;
      const double* ap = A + 2 * ((long)gm * lda + kb + j);
      pa[2 * j] = (gm < m && ok) ? (R)ap[0] : (R)0;
      pa[2 * j + 1] = (gm < m && ok) ? (R)ap[1] : (R)0;
#pragma unroll
      for (int h = 0; h < BROWS; ++h) {
        const int gn = tile_n * TN + srow + 64 * h;
        const R* bp = B + NB * ((long)gn * ldb + kb + j);
        pb[h][NB * j] = (gn < n && ok) ? bp[0] : (R)0;
        if (BC) pb[h][NB * j + 1] = (gn < n && ok) ? bp[1] : (R)0;
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      sA[buf][2 * (srow * LDK + skq + j)] = pa[2 * j];
      sA[buf][2 * (srow * LDK + skq + j) + 1] = pa[2 * j + 1];
#pragma unroll
      for (int h = 0; h < BROWS; ++h) {
        const int row = srow + 64 * h;
        sB[buf][NB * (row * LDK + skq + j)] = pb[h][NB * j];
        if (BC) sB[buf][NB * (row * LDK + skq + j) + 1] = pb[h][NB * j + 1];
      }
    }
  };

  acc_t cr[2][NTW], ci[2][NTW];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NTW; ++b) {
      cr[a][b] = acc_t{0, 0, 0, 0};
      ci[a][b] = acc_t{0, 0, 0, 0};
    }

  if (kbeg < kend) {
    gload(kbeg);
    sstore(0);
  }
  __syncthreads();
  for (int k0 = kbeg, buf = 0; k0 < kend; k0 += KS, buf ^= 1) {
    const bool more = k0 + KS < kend;
    if (more) gload(k0 + KS);
    const R* a_s = sA[buf];
    const R* b_s = sB[buf];
#pragma unroll
    for (int kk = 0; kk < KS; kk += 16) {
      R ar[2][4], ai[2][4], br[NTW][4], bi[NTW][4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int kx = kk + 4 * g + s;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int ra = wm + t * 16 + r16;
          ar[t][s] = a_s[2 * (ra * LDK + kx)];
          ai[t][s] = a_s[2 * (ra * LDK + kx) + 1];
        }
#pragma unroll
        for (int t = 0; t < NTW; ++t) {
          const int rb = wn + t * 16 + r16;
          br[t][s] = b_s[NB * (rb * LDK + kx)];
          bi[t][s] = BC ? b_s[NB * (rb * LDK + kx) + 1] : (R)0;
        }
      }
      // independent accumulators interleaved: 8 (or 4) chains per k-step
#pragma unroll
      for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < NTW; ++b) {
            cr[a][b] = mfma16x16x4(ar[a][s], br[b][s], cr[a][b]);
            ci[a][b] = mfma16x16x4(ai[a][s], br[b][s], ci[a][b]);
          }
        if (BC) {
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < NTW; ++b) {
              cr[a][b] = mfma16x16x4(-ai[a][s], bi[b][s], cr[a][b]);
              ci[a][b] = mfma16x16x4(ar[a][s], bi[b][s], ci[a][b]);
            }
        }
      }
    }
    // the other buffer was last read in the previous slice, before the barrier that ended it
    if (more) sstore(buf ^ 1);
    __syncthreads();
  }

  const int m0 = tile_m * kTile + wm, n0 = tile_n * TN + wn;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NTW; ++b) {
      const int col = n0 + b * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + a * 16 + acc_row<R>(g, r);
        if (row < m && col < n) {
          double2 v;
          v.x = alpha * (double)cr[a][b][r];
          v.y = alpha * (double)ci[a][b][r];
          *reinterpret_cast<double2*>(C + 2 * ((long)row * ldc + col)) = v;
        }
      }
    }
}

// C[row][col] = alpha * sum_z P[z][row][col]   (fixed order over the splits)
__global__ void k_split_reduce(int m, int n, int splits, double alpha,
                               const double2* __restrict__ P, double2* __restrict__ C,
                               long ldc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  double re = 0.0, im = 0.0;
  for (int z = 0; z < splits; ++z) {
    const double2 v = P[(long)z * m * n + e];
    re += v.x;
    im += v.y;
  }
  const long row = e / n, col = e % n;
  C[row * ldc + col] = double2{alpha * re, alpha * im};
}

}  // namespace

void launch_split_reduce(hipStream_t s, int m, int n, int splits, double alpha,
                         const double* split_ws, double* C, long ldc) {
  const long tot = (long)m * n;
  hipLaunchKernelGGL(k_split_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, n,
                     splits, alpha, reinterpret_cast<const double2*>(split_ws),
                     reinterpret_cast<double2*>(C), ldc);
  AFQ_HIP(hipGetLastError());
}

int gemm_splits(int m, int n, int k) {
  const int tiles = ((m + kTile - 1) / kTile) * ((n + kTile - 1) / kTile);
  int s = (512 + tiles - 1) / tiles;
  s = std::min(s, std::max(1, k / 128));
  return std::max(1, std::min(s, kMaxSplits));
}

int gemm_tile_configs(int precision, bool b_complex) {
  // FP32: 16- or 32-wide k slices, and for real B also 128-wide column tiles; FP64: one tiling
  // (a 64 x 128 FP64 tile or 32-wide FP64 slices would not fit the register or LDS budget)
  if (precision == AFQ_FP64 || precision == AFQ_BF16X3) return 1;
  return b_complex ? 2 : 4;
}

void launch_gemm_nt_z(hipStream_t s, int precision, int m, int n, int k, double alpha,
                      const double* A, long lda, const void* B, long ldb, bool b_complex,
                      double* C, long ldc, double* split_ws, long split_ws_elems, int cfg) {
  if (m <= 0 || n <= 0) return;
  int splits = split_ws ? gemm_splits(m, n, k) : 1;
  if ((long)splits * m * n > split_ws_elems) splits = 1;
  int kchunk = (k + splits - 1) / splits;
  kchunk = (kchunk + 31) & ~31;   // k chunks on whole staged slices (16 or 32 wide)
  splits = (k + kchunk - 1) / kchunk;
  if (splits < 1) splits = 1;
  // 64 x 64 tiles and 16-wide k slices everywhere: the vHS grid (512 x 4096 at c2) then has 512
  // workgroups, two per CU; in the FP32 form 64 x 128 tiles with 32-wide slices (one workgroup
  // per CU) measured 63 us against 56 us, and the FP32 vbias 43 us with 32-wide slices against
  // 36 us
  // cfg (FP32 only; < 0 or out of range: 0): bit 0 = 32-wide k slices, bit 1 = 128-wide column
  // tiles (real B). Every tiling sums k in the same order, so the choice changes no result.
  if (cfg < 0 || cfg >= gemm_tile_configs(precision, b_complex)) cfg = 0;
  const int tn = (cfg & 2) ? 128 : 64;
  // real B (vHS, the large Cholesky tensor): column tiles fastest (XCD-local B)
  const int nmajor = b_complex ? 0 : 1;
  const unsigned mt = (m + kTile - 1) / kTile, nt = (n + tn - 1) / tn;
  dim3 grid(nmajor ? nt : mt, nmajor ? mt : nt, splits);
  double* out = splits > 1 ? split_ws : C;
  const long ldo = splits > 1 ? n : ldc;
  const long sstride = splits > 1 ? 2L * m * n : 0;
  const double a = splits > 1 ? 1.0 : alpha;
  if (precision == AFQ_FP64) {
    if (b_complex)
      hipLaunchKernelGGL((k_gemm_nt<double, true, 16, 64>), grid, dim3(256), 0, s, m, n, k, kchunk,
                         a, A, lda, (const double*)B, ldb, out, ldo, sstride, nmajor);
    else
      hipLaunchKernelGGL((k_gemm_nt<double, false, 16, 64>), grid, dim3(256), 0, s, m, n, k,
                         kchunk, a, A, lda, (const double*)B, ldb, out, ldo, sstride, nmajor);
  } else {
    const float* Bf = (const float*)B;
    if (b_complex) {
      if (cfg & 1)
        hipLaunchKernelGGL((k_gemm_nt<float, true, 32, 64>), grid, dim3(256), 0, s, m, n, k,
                           kchunk, a, A, lda, Bf, ldb, out, ldo, sstride, nmajor);
      else
        hipLaunchKernelGGL((k_gemm_nt<float, true, 16, 64>), grid, dim3(256), 0, s, m, n, k,
                           kchunk, a, A, lda, Bf, ldb, out, ldo, sstride, nmajor);
    } else {
      switch (cfg) {
        case 1:
          hipLaunchKernelGGL((k_gemm_nt<float, false, 32, 64>), grid, dim3(256), 0, s, m, n, k,
                             kchunk, a, A, lda, Bf, ldb, out, ldo, sstride, nmajor);
          break;
        case 2:
          hipLaunchKernelGGL((k_gemm_nt<float, false, 16, 128>), grid, dim3(256), 0, s, m, n, k,
                             kchunk, a, A, lda, Bf, ldb, out, ldo, sstride, nmajor);
          break;
        case 3:
          hipLaunchKernelGGL((k_gemm_nt<float, false, 32, 128>), grid, dim3(256), 0, s, m, n, k,
                             kchunk, a, A, lda, Bf, ldb, out, ldo, sstride, nmajor);
          break;
        default:
          hipLaunchKernelGGL((k_gemm_nt<float, false, 16, 64>), grid, dim3(256), 0, s, m, n, k,
                             kchunk, a, A, lda, Bf, ldb, out, ldo, sstride, nmajor);
      }
    }
  }
  AFQ_HIP(hipGetLastError());
  if (splits > 1) launch_split_reduce(s, m, n, splits, alpha, split_ws, C, ldc);
}

}  // namespace afq
