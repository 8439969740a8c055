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

// A operand fragment of one 16-wide k-chunk: A[row][k0 + 4g + s], s < 4
__device__ inline void load_a(const cd* __restrict__ A, long lda, int M, int row, int k0, int g,
                              double (&ar)[4], double (&ai)[4]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int kx = k0 + 4 * g + s;
    const bool ok = row < M && kx < M;
    const cd a = ok ? A[(long)row * lda + kx] : cmk(0.0, 0.0);
    ar[s] = a.re;
    ai[s] = a.im;
  }
}

// acc(tiles of this wave) = A[rows][0..M) * sB[0..M)[cols]; A in LDS transposed (A_LDS:
// A[row][k] at A[k * lda + row], zero padded to MB) or in memory (prefetched one k-chunk ahead)
template <int TPW, bool A_LDS>
__device__ inline void mm_tiles(int M, int MB, int NB, int ldb, const cd* __restrict__ A,
                                long lda, const cd* sB, int ntw, const int* tile_ids, d4* accr,
                                d4* acci) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, r16 = lane & 15;
  const int ntc = NB / 16;
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
      double ar[4], ai[4], nr[4], ni[4];
      if (A_LDS) {
        // rows and columns beyond M are zero in LDS
      } else {
        load_a(A, lda, M, row, 0, g, ar, ai);
      }
      for (int k0 = 0; k0 < MB; k0 += 16) {
        if (A_LDS) {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const cd a = A[(long)(k0 + 4 * g + s) * lda + row];
            ar[s] = a.re;
            ai[s] = a.im;
          }
        } else if (k0 + 16 < MB) {
          load_a(A, lda, M, row, k0 + 16, g, nr, ni);
        }
        double br[4], bi[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const cd b = sB[(k0 + 4 * g + s) * ldb + tj * 16 + r16];
          br[s] = b.re;
          bi[s] = b.im;
        }
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

template <int TPW, bool V_LDS>
__global__ __launch_bounds__(64 * kWaves) void k_propagate(int M, int N, cd* __restrict__ Phi,
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
