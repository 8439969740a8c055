// Walker propagation Phi <- P1 exp(iV) P1 Phi with exp(iV) as its order-n Taylor series
// (det_ops::detail::propagate_impl + apply_expM, propagate.hpp:41-70, 241-298):
//   S = P1 Phi;  T = S;  for k = 1..order: T = (i/k) V_w T;  S += T;   Phi = P1 S
//
// One workgroup (4 wavefronts) per walker. V_w is staged in LDS once (batched loads) when it
// fits, so the Taylor terms read it order times from LDS rather than from memory; P1 (shared by
// all walkers, L2-resident) is read from memory with the next k-chunk's operands in flight
// during the current chunk's MFMAs. The current Taylor term T lives in LDS as the MFMA B
// operand ([M][NB] complex, NB = N rounded up to 16); the running sum S and each new term stay in
// registers in the MFMA accumulator layout, each wavefront owning a fixed set of 16x16 output
// tiles. V_w is stored transposed in LDS, so that the 8 lanes of an LDS phase read 8 adjacent
// rows of V_w as 128 contiguous bytes, like the B operand. Complex products take 4 FP64 MFMA
// 16x16x4 per tile and k-step.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kWaves = 4;

// acc(tiles of this wave) = A[rows][0..M) * sB[0..M)[cols]; A in LDS transposed (A_LDS:
// A[row][k] at A[k * lda + row], zero padded to MB) or in memory (prefetched one k-chunk ahead)
template <int TPW, bool A_LDS>
__device__ inline void mm_tiles(int M, int MB, int NB, int ldb, const cd* __restrict__ A,
                                long lda, const cd* sB, int ntw, const int* tile_ids, d4* accr,
                                d4* acci) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, r16 = lane & 15;
  const int ntc = NB / 16;
  // strides of the four k values a lane group reads per chunk (no multiplies in the k loop)
  const int lb1 = ldb, lb2 = 2 * ldb, lb3 = 3 * ldb;
  const long la1 = lda, la2 = 2 * lda, la3 = 3 * lda;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    accr[q] = d4{0, 0, 0, 0};
    acci[q] = d4{0, 0, 0, 0};
  }
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    if (q < ntw) {
      const int ti = tile_ids[q] / ntc, tj = tile_ids[q] % ntc;
      const int row = ti * 16 + r16;
      // A fragment of chunk k0: A[row][k0 + 4g + s]; in LDS A is transposed (A[k][row])
      const cd* ap = A_LDS ? A + (long)(4 * g) * lda + row : A + (long)row * lda + 4 * g;
      const bool arow = row < M;
      const cd* bp = sB + (4 * g) * ldb + tj * 16 + r16;
      double ar[4], ai[4], nr[4], ni[4];
      auto load_mem = [&](const cd* p, int k0, double (&xr)[4], double (&xi)[4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const bool ok = arow && k0 + 4 * g + s < M;
          const cd x = ok ? p[s] : cmk(0.0, 0.0);
          xr[s] = x.re;
          xi[s] = x.im;
        }
      };
      if (!A_LDS) load_mem(ap, 0, ar, ai);
      for (int k0 = 0; k0 < MB; k0 += 16) {
        if (A_LDS) {
          const cd x0 = ap[0], x1 = ap[la1], x2 = ap[la2], x3 = ap[la3];
          ar[0] = x0.re; ai[0] = x0.im; ar[1] = x1.re; ai[1] = x1.im;
          ar[2] = x2.re; ai[2] = x2.im; ar[3] = x3.re; ai[3] = x3.im;
          ap += 16 * lda;
        } else if (k0 + 16 < MB) {
          load_mem(ap + 16, k0 + 16, nr, ni);
        }
        const cd y0 = bp[0], y1 = bp[lb1], y2 = bp[lb2], y3 = bp[lb3];
        bp += 16 * ldb;
        const double br[4] = {y0.re, y1.re, y2.re, y3.re};
        const double bi[4] = {y0.im, y1.im, y2.im, y3.im};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          accr[q] = mfma16x16x4(ar[s], br[s], accr[q]);
          acci[q] = mfma16x16x4(ar[s], bi[s], acci[q]);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          accr[q] = mfma16x16x4(-ai[s], bi[s], accr[q]);
          acci[q] = mfma16x16x4(ai[s], br[s], acci[q]);
        }
        if (!A_LDS) {
          ap += 16;
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            ar[s] = nr[s];
            ai[s] = ni[s];
          }
        }
      }
    }
  }
}

template <int TPW>
__device__ inline void store_tiles(cd* sB, int NB, int ldb, int ntw, const int* tile_ids,
                                   const d4* xr, const d4* xi) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, r16 = lane & 15;
  const int ntc = NB / 16;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    if (q < ntw) {
      const int ti = tile_ids[q] / ntc, tj = tile_ids[q] % ntc;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sB[(ti * 16 + acc_row<double>(g, r)) * ldb + tj * 16 + r16] = cmk(xr[q][r], xi[q][r]);
    }
  }
}

// LDS (80 KB per workgroup with V_w staged) allows two workgroups per CU, i.e. two wavefronts
// per SIMD, so the register budget is that of two wavefronts (no accumulators parked in AGPRs)
template <int TPW, bool V_LDS>
__global__ __launch_bounds__(64 * kWaves, 2) void k_propagate(int M, int N, cd* __restrict__ Phi,
                                                          long ldw, const cd* __restrict__ P1,
                                                          const cd* __restrict__ V, int order) {
  extern __shared__ cd smem[];   // [MB][NB] current term, then (V_LDS) [MB][MB] V_w transposed
  const int w = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int MB = (M + 15) & ~15, NB = (N + 15) & ~15;
  const int ldb = NB, ldv = MB;
  cd* sB = smem;
  cd* sV = smem + MB * ldb;
  const int ntc = NB / 16, ntiles = (MB / 16) * ntc;
  int tile_ids[TPW];
  int ntw = 0;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = wave + q * kWaves;
    tile_ids[q] = t;
    if (t < ntiles) ntw = q + 1;
  }

  cd* P = Phi + (long)w * ldw;
  const cd* Vw = V + (long)w * M * M;
  if (V_LDS) {
    // V_w^T [MB][MB], zero outside M x M; 8 loads per thread in flight per batch
    const int tot = MB * MB;
    for (int e0 = 0; e0 < tot; e0 += 8 * 64 * kWaves) {
      cd v[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int e = e0 + b * 64 * kWaves + threadIdx.x;
        const int i = e / MB, j = e % MB;
        v[b] = (e < tot && i < M && j < M) ? Vw[(long)i * M + j] : cmk(0.0, 0.0);
      }
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int e = e0 + b * 64 * kWaves + threadIdx.x;
        if (e < tot) sV[(e % MB) * ldv + e / MB] = v[b];
      }
    }
  }
  for (int e = threadIdx.x; e < MB * NB; e += 64 * kWaves) {
    const int i = e / NB, b = e % NB;
    sB[i * ldb + b] = (i < M && b < N) ? P[i * N + b] : cmk(0.0, 0.0);
  }
  __syncthreads();

  d4 sr[TPW], si[TPW], tr[TPW], ti[TPW];
  // S = P1 Phi
  mm_tiles<TPW, false>(M, MB, NB, ldb, P1, M, sB, ntw, tile_ids, sr, si);
  __syncthreads();
  store_tiles<TPW>(sB, NB, ldb, ntw, tile_ids, sr, si);
  __syncthreads();
  for (int k = 1; k <= order; ++k) {
    if (V_LDS)
      mm_tiles<TPW, true>(M, MB, NB, ldb, sV, ldv, sB, ntw, tile_ids, tr, ti);
    else
      mm_tiles<TPW, false>(M, MB, NB, ldb, Vw, M, sB, ntw, tile_ids, tr, ti);
    const double f = 1.0 / (double)k;
    // T <- (i/k) (V T):  re = -f im, im = f re
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const d4 nr = -f * ti[q], ni = f * tr[q];
      tr[q] = nr;
      ti[q] = ni;
      sr[q] += nr;
      si[q] += ni;
    }
    __syncthreads();
    store_tiles<TPW>(sB, NB, ldb, ntw, tile_ids, tr, ti);
    __syncthreads();
  }
  // Phi = P1 S
  store_tiles<TPW>(sB, NB, ldb, ntw, tile_ids, sr, si);
  __syncthreads();
  mm_tiles<TPW, false>(M, MB, NB, ldb, P1, M, sB, ntw, tile_ids, tr, ti);
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    if (q < ntw) {
      const int ti_ = tile_ids[q] / ntc, tj = tile_ids[q] % ntc;
      const int b = tj * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = ti_ * 16 + acc_row<double>(g, r);
        if (i < M && b < N) P[i * N + b] = cmk(tr[q][r], ti[q][r]);
      }
    }
  }
}

}  // namespace

void launch_propagate(hipStream_t s, int nb, int M, int N, double* Phi, long ldw,
                      const double* P1, const double* V, int order) {
  if (nb <= 0 || N <= 0) return;
  const int MB = (M + 15) & ~15, NB = (N + 15) & ~15;
  const int ntiles = (MB / 16) * (NB / 16);
  const int tpw = (ntiles + kWaves - 1) / kWaves;
  AFQ_CHECK(tpw <= 8, AFQ_EUNSUPPORTED, "propagate: walker matrix larger than 128 x 64");
  const size_t lds_b = sizeof(cd) * (size_t)MB * NB;
  const size_t lds_v = sizeof(cd) * (size_t)MB * MB;
  // V_w in LDS when both fit in 80 KB (two workgroups per CU)
  const bool vlds = lds_b + lds_v <= 80 * 1024;
  const size_t lds = vlds ? lds_b + lds_v : lds_b;
  auto args = [&](auto kern) {
    ensure_dynamic_lds((const void*)kern, lds);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64 * kWaves), lds, s, M, N,
                       reinterpret_cast<cd*>(Phi), ldw, reinterpret_cast<const cd*>(P1),
                       reinterpret_cast<const cd*>(V), order);
  };
  if (vlds) {
    if (tpw <= 1) args(k_propagate<1, true>);
    else if (tpw <= 2) args(k_propagate<2, true>);
    else if (tpw <= 4) args(k_propagate<4, true>);
    else args(k_propagate<8, true>);
  } else {
    if (tpw <= 1) args(k_propagate<1, false>);
    else if (tpw <= 2) args(k_propagate<2, false>);
    else if (tpw <= 4) args(k_propagate<4, false>);
    else args(k_propagate<8, false>);
  }
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
