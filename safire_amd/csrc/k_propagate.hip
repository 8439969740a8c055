// Walker propagation Phi <- P1 exp(iV) P1 Phi with exp(iV) as its order-n Taylor series
// (det_ops::detail::propagate_impl + apply_expM, propagate.hpp:41-70, 241-298):
//   S = P1 Phi;  T = S;  for k = 1..order: T = (i/k) V_w T;  S += T;   Phi = P1 S
//
// One workgroup (4 wavefronts) per walker. The current Taylor term T lives in LDS as the MFMA
// B operand ([M][NB] complex, NB = N rounded up to 16); the running sum S and each new term stay
// in registers in the MFMA accumulator layout, each wavefront owning a fixed set of 16x16
// output tiles. V_w and P1 rows are the A operand, read from memory 64 bytes per lane and
// k-chunk. Complex products take 4 FP64 MFMA 16x16x4 per tile and k-step.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kWaves = 4;

// acc(tiles of this wave) = A[rows][0..M) * sB[0..M)[cols]   (A: c128 row-major, lda)
template <int TPW>
__device__ inline void mm_tiles(int M, int MB, int NB, const cd* __restrict__ A, long lda,
                                const cd* sB, int ntw, const int* tile_ids, d4* accr,
                                d4* acci) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, r16 = lane & 15;
  const int ntc = NB / 16;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    accr[q] = d4{0, 0, 0, 0};
    acci[q] = d4{0, 0, 0, 0};
  }
  for (int k0 = 0; k0 < MB; k0 += 16) {
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      if (q < ntw) {
        const int ti = tile_ids[q] / ntc, tj = tile_ids[q] % ntc;
        const int row = ti * 16 + r16;
        double ar[4], ai[4], br[4], bi[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int kx = k0 + 4 * g + s;
          const bool ok = row < M && kx < M;
          const cd a = ok ? A[(long)row * lda + kx] : cmk(0.0, 0.0);
          ar[s] = a.re;
          ai[s] = a.im;
          const cd b = sB[kx * NB + tj * 16 + r16];
          br[s] = b.re;
          bi[s] = b.im;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          accr[q] = mfma16x16x4(ar[s], br[s], accr[q]);
          accr[q] = mfma16x16x4(-ai[s], bi[s], accr[q]);
          acci[q] = mfma16x16x4(ar[s], bi[s], acci[q]);
          acci[q] = mfma16x16x4(ai[s], br[s], acci[q]);
        }
      }
    }
  }
}

template <int TPW>
__device__ inline void store_tiles(cd* sB, int NB, int ntw, const int* tile_ids, const d4* xr,
                                   const d4* xi) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, r16 = lane & 15;
  const int ntc = NB / 16;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    if (q < ntw) {
      const int ti = tile_ids[q] / ntc, tj = tile_ids[q] % ntc;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sB[(ti * 16 + acc_row<double>(g, r)) * NB + tj * 16 + r16] = cmk(xr[q][r], xi[q][r]);
    }
  }
}

template <int TPW>
__global__ __launch_bounds__(64 * kWaves) void k_propagate(int M, int N, cd* __restrict__ Phi,
                                                          long ldw, const cd* __restrict__ P1,
                                                          const cd* __restrict__ V, int order) {
  extern __shared__ cd sB[];   // [MB][NB]
  const int w = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int MB = (M + 15) & ~15, NB = (N + 15) & ~15;
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
  for (int e = threadIdx.x; e < MB * NB; e += 64 * kWaves) {
    const int i = e / NB, b = e % NB;
    sB[e] = (i < M && b < N) ? P[i * N + b] : cmk(0.0, 0.0);
  }
  __syncthreads();

  d4 sr[TPW], si[TPW], tr[TPW], ti[TPW];
  // S = P1 Phi
  mm_tiles<TPW>(M, MB, NB, P1, M, sB, ntw, tile_ids, sr, si);
  __syncthreads();
  store_tiles<TPW>(sB, NB, ntw, tile_ids, sr, si);
  __syncthreads();
  for (int k = 1; k <= order; ++k) {
    mm_tiles<TPW>(M, MB, NB, Vw, M, sB, ntw, tile_ids, tr, ti);
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
    store_tiles<TPW>(sB, NB, ntw, tile_ids, tr, ti);
    __syncthreads();
  }
  // Phi = P1 S
  store_tiles<TPW>(sB, NB, ntw, tile_ids, sr, si);
  __syncthreads();
  mm_tiles<TPW>(M, MB, NB, P1, M, sB, ntw, tile_ids, tr, ti);
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
  const size_t lds = sizeof(cd) * (size_t)MB * NB;
  auto args = [&](auto kern) {
    if (lds > 64 * 1024)
      AFQ_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64 * kWaves), lds, s, M, N,
                       reinterpret_cast<cd*>(Phi), ldw, reinterpret_cast<const cd*>(P1),
                       reinterpret_cast<const cd*>(V), order);
  };
  if (tpw <= 1) args(k_propagate<1>);
  else if (tpw <= 2) args(k_propagate<2>);
  else if (tpw <= 4) args(k_propagate<4>);
  else args(k_propagate<8>);
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
