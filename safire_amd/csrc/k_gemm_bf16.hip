// BF16X3 form of the Cholesky GEMMs C[m][n] = alpha * sum_k A[m][k] B[n][k] (A complex FP64,
// B complex or real, C complex FP64) on v_mfma_f32_16x16x32_bf16:
//   vbias  (Real3IndexFactorization.hpp:350-361): A = G [W][N*M], B = Lnak [K][N*M]  (complex)
//   vHS    (Real3IndexFactorization.hpp:297-313): A = X [W][K],   B = Likn [M*M][K]   (real)
//
// Arithmetic: every real operand x = h + l with h = bf16(x), l = bf16(x - h); a real product is
// h h' + h l' + l h' (three MFMAs, FP32 accumulation). A complex B uses the three-product
// (Gauss) complex multiplication: P1 = Re a Re b, P2 = Im a Im b, P3 = (Re a + Im a)(Re b + Im b),
// Re = P1 - P2, Im = P3 - P1 - P2, so 9 MFMAs per complex 16x16x32 tile; a real B needs 6 (Re a b,
// Im a b). B is split once at set-up into BF16 planes (hi, lo of Re, Im, Re + Im; or hi, lo),
// rows padded with zeros to a multiple of 32; A is split when it is staged.
//
// Tiling: a 256-thread workgroup owns a 64 x 64 tile of C, each wavefront 32 x 32 (2 x 2 MFMA
// tiles); each 32-wide k slice of A and B is staged in LDS as planes of 64 rows padded to 40
// BF16 (80 bytes: the 16 rows of a fragment read fall into distinct bank groups), the next
// slice's loads in flight in registers while the current one is multiplied. k may be split over
// grid.z into a workspace summed by k_split_reduce (fixed order).
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kTile = 64;
constexpr int kKS = 32;          // k slice
constexpr int kLdk = 40;         // staged row stride (bf16)

__device__ inline void split2(double x, __bf16& h, __bf16& l) {
  h = (__bf16)(float)x;
  l = (__bf16)(float)(x - (double)(float)h);
}

__device__ inline f4 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f4 c) {
  c = mfma_bf16(ah, bh, c);
  c = mfma_bf16(ah, bl, c);
  return mfma_bf16(al, bh, c);
}

template <bool BC>
__global__ __launch_bounds__(256) void k_gemm_bf16(int m, int n, int k, int kchunk, double alpha,
                                                   const double* __restrict__ A, long lda,
                                                   const __bf16* __restrict__ Bp, long ldb,
                                                   long bplane, double* __restrict__ C,
                                                   long ldc, long split_stride) {
  constexpr int PA = BC ? 6 : 4;     // A planes: hi, lo of Re, Im (and Re + Im)
  constexpr int PB = BC ? 6 : 2;     // B planes: hi, lo of Re, Im, Re + Im; or hi, lo
  __shared__ __attribute__((aligned(16))) __bf16 sA[PA][kTile * kLdk];
  __shared__ __attribute__((aligned(16))) __bf16 sB[PB][kTile * kLdk];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int tile_m = blockIdx.x, tile_n = blockIdx.y;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(k, kbeg + kchunk);
  C += blockIdx.z * split_stride;

  // staging share: row srow of the A and B tiles, k = k0 + 8 sq .. k0 + 8 sq + 7
  const int srow = tid >> 2, sq = (tid & 3) * 8;
  const int gm = tile_m * kTile + srow, gn = tile_n * kTile + srow;
  double pa[16];
  bf16x8 pb[PB];
  auto gload = [&](int k0) {
    const int kb = k0 + sq;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const bool ok = gm < m && kb + j < kend;
      pa[2 * j] = ok ? A[2 * ((long)gm * lda + kb + j)] : 0.0;
      pa[2 * j + 1] = ok ? A[2 * ((long)gm * lda + kb + j) + 1] : 0.0;
    }
    // B rows are padded to whole slices with zeros; the last partial slice of a split is cut
    // by kend, which is a multiple of 32 except at k itself (past which B holds zeros)
#pragma unroll
    for (int p = 0; p < PB; ++p)
      pb[p] = gn < n ? *reinterpret_cast<const bf16x8*>(Bp + p * bplane + (long)gn * ldb + kb)
                     : bf16x8{};
  };
  auto sstore = [&]() {
    bf16x8 h[PA];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      __bf16 hh, ll;
      split2(pa[2 * j], hh, ll);
      h[0][j] = hh;
      h[1][j] = ll;
      split2(pa[2 * j + 1], hh, ll);
      h[2][j] = hh;
      h[3][j] = ll;
      if constexpr (BC) {
        split2(pa[2 * j] + pa[2 * j + 1], hh, ll);
        h[4][j] = hh;
        h[5][j] = ll;
      }
    }
#pragma unroll
    for (int p = 0; p < PA; ++p)
      *reinterpret_cast<bf16x8*>(&sA[p][srow * kLdk + sq]) = h[p];
#pragma unroll
    for (int p = 0; p < PB; ++p) *reinterpret_cast<bf16x8*>(&sB[p][srow * kLdk + sq]) = pb[p];
  };

  // complex B: P1, P2, P3 per tile; real B: Re, Im per tile
  constexpr int NACC = BC ? 3 : 2;
  f4 acc[2][2][NACC];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int j = 0; j < NACC; ++j) acc[a][b][j] = f4{0.f, 0.f, 0.f, 0.f};

  if (kbeg < kend) gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += kKS) {
    __syncthreads();   // the previous slice has been read by every wavefront
    sstore();
    __syncthreads();
    if (k0 + kKS < kend) gload(k0 + kKS);
    bf16x8 fa[2][PA], fb[2][PB];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int p = 0; p < PA; ++p)
        fa[t][p] = *reinterpret_cast<const bf16x8*>(&sA[p][(wm + 16 * t + r16) * kLdk + 8 * g]);
#pragma unroll
      for (int p = 0; p < PB; ++p)
        fb[t][p] = *reinterpret_cast<const bf16x8*>(&sB[p][(wn + 16 * t + r16) * kLdk + 8 * g]);
    }
#pragma unroll
    for (int j = 0; j < NACC; ++j)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          if constexpr (BC)   // P1: Re Re, P2: Im Im, P3: sum sum
            acc[a][b][j] = mfma3(fa[a][2 * j], fa[a][2 * j + 1], fb[b][2 * j],
                                 fb[b][2 * j + 1], acc[a][b][j]);
          else      // Re a b, Im a b
            acc[a][b][j] = mfma3(fa[a][2 * j], fa[a][2 * j + 1], fb[b][0], fb[b][1],
                                 acc[a][b][j]);
        }
  }

  const int m0 = tile_m * kTile + wm, n0 = tile_n * kTile + wn;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = n0 + b * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + a * 16 + acc_row<float>(g, r);
        if (row < m && col < n) {
          double re, im;
          if constexpr (BC) {
            const double p1 = acc[a][b][0][r], p2 = acc[a][b][1][r];
            re = p1 - p2;
            im = (double)acc[a][b][NACC - 1][r] - p1 - p2;
          } else {
            re = acc[a][b][0][r];
            im = acc[a][b][1][r];
          }
          double2 v;
          v.x = alpha * re;
          v.y = alpha * im;
          *reinterpret_cast<double2*>(C + 2 * ((long)row * ldc + col)) = v;
        }
      }
    }
}

// planes of a [rows][cols] matrix (complex interleaved or real), rows padded to colsp with zeros:
// complex: hi, lo of Re, Im, Re + Im; real: hi, lo. One thread per padded element.
__global__ void k_split_planes_bf16(long rows, int cols, int colsp, int complex_src,
                                    const double* __restrict__ src, __bf16* __restrict__ dst) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * colsp) return;
  const long r = e / colsp;
  const int c = (int)(e % colsp);
  const long plane = rows * colsp;
  __bf16 h, l;
  if (complex_src) {
    double xr = 0.0, xi = 0.0;
    if (c < cols) {
      xr = src[2 * (r * cols + c)];
      xi = src[2 * (r * cols + c) + 1];
    }
    split2(xr, h, l);
    dst[e] = h;
    dst[plane + e] = l;
    split2(xi, h, l);
    dst[2 * plane + e] = h;
    dst[3 * plane + e] = l;
    split2(xr + xi, h, l);
    dst[4 * plane + e] = h;
    dst[5 * plane + e] = l;
  } else {
    split2(c < cols ? src[r * cols + c] : 0.0, h, l);
    dst[e] = h;
    dst[plane + e] = l;
  }
}

}  // namespace

int bf16_row(int cols) { return (cols + kKS - 1) / kKS * kKS; }

void split_planes_bf16(hipStream_t s, long rows, int cols, bool complex_src, const double* src,
                       __bf16* dst) {
  const int colsp = bf16_row(cols);
  const long tot = rows * colsp;
  hipLaunchKernelGGL(k_split_planes_bf16, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                     rows, cols, colsp, complex_src ? 1 : 0, src, dst);
  AFQ_HIP(hipGetLastError());
}

void launch_gemm_bf16(hipStream_t s, int m, int n, int k, double alpha, const double* A, long lda,
                      const __bf16* Bp, bool b_complex, double* C, long ldc, double* split_ws,
                      long split_ws_elems) {
  if (m <= 0 || n <= 0) return;
  const int kp = bf16_row(k);
  int splits = split_ws ? gemm_splits(m, n, k) : 1;
  if ((long)splits * m * n > split_ws_elems) splits = 1;
  int kchunk = (k + splits - 1) / splits;
  kchunk = (kchunk + kKS - 1) / kKS * kKS;   // splits on whole slices
  splits = (k + kchunk - 1) / kchunk;
  if (splits < 1) splits = 1;
  dim3 grid((m + kTile - 1) / kTile, (n + kTile - 1) / kTile, splits);
  double* out = splits > 1 ? split_ws : C;
  const long ldo = splits > 1 ? n : ldc;
  const long sstride = splits > 1 ? 2L * m * n : 0;
  const double a = splits > 1 ? 1.0 : alpha;
  const long bplane = (long)n * kp;
  if (b_complex)
    hipLaunchKernelGGL(k_gemm_bf16<true>, grid, dim3(256), 0, s, m, n, k, kchunk, a, A, lda, Bp,
                       (long)kp, bplane, out, ldo, sstride);
  else
    hipLaunchKernelGGL(k_gemm_bf16<false>, grid, dim3(256), 0, s, m, n, k, kchunk, a, A, lda, Bp,
                       (long)kp, bplane, out, ldo, sstride);
  AFQ_HIP(hipGetLastError());
  if (splits > 1) launch_split_reduce(s, m, n, splits, alpha, split_ws, C, ldc);
}

}  // namespace afq
