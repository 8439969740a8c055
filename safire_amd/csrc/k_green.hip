// Batched mixed density matrix and log-overlap of closed-shell walkers, one wavefront per walker.
//
// det_ops::MixedDensityMatrix / detail::log_overlap_impl (density_matrix.hpp:113-156, 535-601):
//   T = PsiT Phi_w                  [N,N]
//   partial pivoting (LAPACK getrf: pivot = first max of |re|+|im| in the column)
//   ov += sum_i log(+-U_ii + DBL_MIN), the sign flipped where rows were swapped, imaginary part
//        wrapped to [-pi,pi]        (log_determinant_from_getrf, determinants_impl.hpp:32-75)
//   T^-1, all zero when a pivot is exactly 0 (getri_or_zero, nda_functions.hpp:422-438)
//   G = T^-T Phi_w^T                [N,M] compact Green's function
// Optional epilogue: E1 = E0 + scl * sum_{a,j} G[a][j] haj[a][j] (Real3IndexFactorization.hpp:505-508).
//
// Both products run on FP64 MFMA 16x16x4 tiles. When Phi_w, PsiT and haj fit in 96 KB of LDS
// they are staged there first with batched loads (every load of a batch in flight before the
// first LDS write), so the kernel waits on memory once rather than once per chunk; otherwise the
// operands are read straight from memory (64 bytes per lane and k-chunk).
//
// The inverse comes from Gauss-Jordan elimination of [T | I] with partial pivoting: its pivots
// are getrf's (the rows at and below step j see the same eliminations), so the determinant
// convention is unchanged. Lane c holds column c of [T | I] (2N <= 64 columns) in registers; each
// step broadcasts the pivot column through LDS. One 64-thread workgroup per walker: no
// workgroup barrier, and W workgroups spread over all CUs.
#include "afq_internal.hpp"

#include <cfloat>

namespace afq {

namespace {

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double readlane_d(double v, int lane) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// dst[e] = src[e] for e < n; the B loads of a batch are all in flight before the first store
template <int B>
__device__ inline void stage_batch(cd* dst, const cd* __restrict__ src, int n, int lane) {
  for (int e0 = 0; e0 < n; e0 += 64 * B) {
    double2 v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int e = min(e0 + 64 * b + lane, n - 1);
      v[b] = reinterpret_cast<const double2*>(src)[e];
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int e = e0 + 64 * b + lane;
      if (e < n) reinterpret_cast<double2*>(dst)[e] = v[b];
    }
  }
}

template <int NMAX, bool STAGE>
__global__ __launch_bounds__(64) void k_green(int M, int N, const cd* __restrict__ PsiT,
                                              const cd* __restrict__ Phi, long ldw,
                                              cd* __restrict__ G, cd* __restrict__ ov,
                                              int ov_mode, double ov_scale, cd ov_add,
                                              const cd* __restrict__ haj, double E0,
                                              double e1_scale, cd* __restrict__ E, long ldE) {
  constexpr int NT = (NMAX + 15) / 16;     // 16x16 tiles along N
  constexpr int NB = 16 * NT;
  __shared__ cd sT[NB * NB];               // T, then T^-1, row-major
  __shared__ cd sCol[2][NMAX];             // pivot column broadcast (double-buffered)
  extern __shared__ cd sDyn[];             // STAGE: Phi_w [M][N], PsiT [N][M], haj [N][M]
  const int lane = threadIdx.x;
  const int g = lane >> 4, r16 = lane & 15;
  const int w = blockIdx.x;
  const cd* Pg = Phi + (long)w * ldw;
  const bool want_e1 = E != nullptr && haj != nullptr;
  if (STAGE) {
    stage_batch<16>(sDyn, Pg, M * N, lane);
    stage_batch<16>(sDyn + M * N, PsiT, M * N, lane);
    if (want_e1) stage_batch<16>(sDyn + 2 * M * N, haj, M * N, lane);
    wave_sync();
  }
  const cd* P = STAGE ? sDyn : Pg;
  const cd* Ps = STAGE ? sDyn + M * N : PsiT;
  const cd* Hj = (STAGE && want_e1) ? sDyn + 2 * M * N : haj;

  // ---- T[a][c] = sum_i PsiT[a][i] Phi[i][c]  (MFMA; k = orbital i) ----
  {
    d4 tr[NT][NT], ti[NT][NT];
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        tr[a][b] = d4{0, 0, 0, 0};
        ti[a][b] = d4{0, 0, 0, 0};
      }
    for (int k0 = 0; k0 < M; k0 += 16) {
      double ar[NT][4], ai[NT][4], br[NT][4], bi[NT][4];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int a = t * 16 + r16;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int k = k0 + 4 * g + s;
          const cd x = (a < N && k < M) ? Ps[(long)a * M + k] : cmk(0, 0);
          ar[t][s] = x.re;
          ai[t][s] = x.im;
          const int c = t * 16 + r16;
          const cd y = (c < N && k < M) ? P[(long)k * N + c] : cmk(0, 0);
          br[t][s] = y.re;
          bi[t][s] = y.im;
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
          for (int b = 0; b < NT; ++b) {
            tr[a][b] = mfma16x16x4(ar[a][s], br[b][s], tr[a][b]);
            ti[a][b] = mfma16x16x4(ar[a][s], bi[b][s], ti[a][b]);
          }
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
          for (int b = 0; b < NT; ++b) {
            tr[a][b] = mfma16x16x4(-ai[a][s], bi[b][s], tr[a][b]);
            ti[a][b] = mfma16x16x4(ai[a][s], br[b][s], ti[a][b]);
          }
      }
    }
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
      for (int b = 0; b < NT; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          sT[(a * 16 + acc_row<double>(g, r)) * NB + b * 16 + r16] = cmk(tr[a][b][r], ti[a][b][r]);
  }
  wave_sync();

  // ---- Gauss-Jordan on [T | I]: lane c < 2N holds column c ----
  cd x[NMAX];
#pragma unroll
  for (int r = 0; r < NMAX; ++r) {
    cd v = cmk(0.0, 0.0);
    if (r < N) {
      if (lane < N) v = sT[r * NB + lane];
      else if (lane - N == r) v = cmk(1.0, 0.0);
    }
    x[r] = v;
  }
  cd upiv = cmk(1.0, 0.0);     // lane j keeps +-U_jj; the logs are taken in parallel afterwards
  bool singular = false;
  if constexpr (NMAX <= 16) {
    // Step j is unrolled, so column j is a register of lane j: lane j finds the pivot in its own
    // column (izamax: first maximum of |re|+|im| from row j down) and readlane broadcasts it; the
    // row swap selects over a static row range; lane j publishes its swapped column to LDS, from
    // which every lane reads the multipliers at fixed addresses (one LDS round trip per step,
    // double-buffered so that the next step's write cannot race this step's reads).
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      bool go = j < N;
      double best = cabs1(x[j]);
      int bi = j;
#pragma unroll
      for (int r = j + 1; r < NMAX; ++r) {
        if (r < N) {
          const double v = cabs1(x[r]);
          if (v > best) {
            best = v;
            bi = r;
          }
        }
      }
      const int p = __builtin_amdgcn_readlane(bi, j);
      cd xp = x[j];
#pragma unroll
      for (int r = j + 1; r < NMAX; ++r)
        if (r == p) xp = x[r];
      const cd piv = cmk(readlane_d(xp.re, j), readlane_d(xp.im, j));
      if (go && lane == j) upiv = (p == j) ? piv : -piv;   // the sign flips where the pivot row moved
      if (go && cis0(piv)) {
        singular = true;   // getrf records the zero pivot and goes on; getri_or_zero zeroes T^-1
        go = false;
      }
      if (go) {
        if (p != j) {
#pragma unroll
          for (int r = j + 1; r < NMAX; ++r)
            if (r == p) x[r] = x[j];
          x[j] = xp;
        }
        cd* col = sCol[j & 1];
        if (lane == j)
#pragma unroll
          for (int r = 0; r < NMAX; ++r)
            if (r < N) col[r] = x[r];
        wave_sync();
        const cd xjn = x[j] * cdiv(cmk(1.0, 0.0), piv);
#pragma unroll
        for (int r = 0; r < NMAX; ++r) {
          if (r < N && r != j) {
            const cd f = col[r];
            if (!cis0(f)) x[r] -= f * xjn;
          }
        }
        x[j] = xjn;
      }
    }
  } else {
    // N > 16: the unrolled form exceeds the unroll budget; step j stays a loop variable and
    // the pivot search is a wavefront reduction over the column broadcast through LDS
    for (int j = 0; j < N; ++j) {
      wave_sync();
      if (lane == j)
#pragma unroll
        for (int r = 0; r < NMAX; ++r)
          if (r < N) sCol[0][r] = x[r];
      wave_sync();
      double v = -1.0;
      int idx = 0x7fffffff;
      if (lane >= j && lane < N) {
        v = cabs1(sCol[0][lane]);
        idx = lane;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov_ = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(idx, off, 64);
        if (ov_ > v || (ov_ == v && oi < idx)) {
          v = ov_;
          idx = oi;
        }
      }
      const int p = idx;
      const cd piv = sCol[0][p];
      if (lane == j) upiv = (p == j) ? piv : -piv;   // the sign flips where the pivot row moved
      if (cis0(piv)) {
        singular = true;   // getrf records the zero pivot and goes on; getri_or_zero zeroes T^-1
        continue;
      }
      if (p != j) {
        cd xj = cmk(0, 0), xp = cmk(0, 0);
#pragma unroll
        for (int r = 0; r < NMAX; ++r) {
          if (r == j) xj = x[r];
          if (r == p) xp = x[r];
        }
#pragma unroll
        for (int r = 0; r < NMAX; ++r) {
          if (r == j) x[r] = xp;
          if (r == p) x[r] = xj;
        }
      }
      const cd rinv = cdiv(cmk(1.0, 0.0), piv);
      cd xjn = cmk(0, 0);
#pragma unroll
      for (int r = 0; r < NMAX; ++r)
        if (r == j) xjn = x[r] * rinv;
#pragma unroll
      for (int r = 0; r < NMAX; ++r) {
        if (r < N) {
          const int rr = (r == j) ? p : ((r == p) ? j : r);   // row r's column-j entry, pre-swap
          const cd f = sCol[0][rr];
          if (r == j) x[r] = xjn;
          else if (!cis0(f)) x[r] -= f * xjn;
        }
      }
    }
  }

  // sum_j log(+-U_jj + DBL_MIN): one log per lane, then a wavefront sum
  cd acc = (lane < N) ? clog_(cmk(upiv.re + DBL_MIN, upiv.im)) : cmk(0.0, 0.0);
  for (int off = 32; off > 0; off >>= 1) {
    acc.re += __shfl_xor(acc.re, off, 64);
    acc.im += __shfl_xor(acc.im, off, 64);
  }
  if (lane == 0) {
    if (ov_mode == 2) acc += ov[w];
    const double pi = 3.14159265358979323846;
    // the reference loops until the value is in range; a non-finite value would never get
    // there, so only finite values are wrapped
    if (!isfinite(acc.im)) {
    } else if (acc.im > pi) {
      while (acc.im > pi) acc.im -= 2 * pi;
    } else if (acc.im < -pi) {
      while (acc.im < -pi) acc.im += 2 * pi;
    }
    if (ov_mode != 0) ov[w] = ov_scale * acc + ov_add;
  }
  if (G == nullptr && E == nullptr) return;

  // ---- T^-1 to LDS as [b][a] (zero padded to NB x NB) ----
  wave_sync();
  for (int e = lane; e < NB * NB; e += 64) sT[e] = cmk(0, 0);
  wave_sync();
  if (lane >= N && lane < 2 * N && !singular) {
    const int a = lane - N;
#pragma unroll
    for (int r = 0; r < NMAX; ++r)
      if (r < N) sT[r * NB + a] = x[r];
  }
  wave_sync();

  // ---- G[a][i] = sum_b Tinv[b][a] Phi[i][b]  (MFMA; rows a, cols i, k = b) ----
  cd e1 = cmk(0.0, 0.0);
  cd* Gw = G ? G + (long)w * N * M : nullptr;
  for (int i0 = 0; i0 < M; i0 += 16) {
    d4 gr[NT], gi[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      gr[t] = d4{0, 0, 0, 0};
      gi[t] = d4{0, 0, 0, 0};
    }
    const int i = i0 + r16;
#pragma unroll
    for (int b0 = 0; b0 < NB; b0 += 16) {
      double ar[NT][4], ai[NT][4], br[4], bi[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int b = b0 + 4 * g + s;
        const cd y = (i < M && b < N) ? P[(long)i * N + b] : cmk(0, 0);
        br[s] = y.re;
        bi[s] = y.im;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const cd z = sT[b * NB + t * 16 + r16];   // Tinv^T[a][b] = Tinv[b][a]
          ar[t][s] = z.re;
          ai[t][s] = z.im;
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          gr[t] = mfma16x16x4(ar[t][s], br[s], gr[t]);
          gi[t] = mfma16x16x4(ar[t][s], bi[s], gi[t]);
          gr[t] = mfma16x16x4(-ai[t][s], bi[s], gr[t]);
          gi[t] = mfma16x16x4(ai[t][s], br[s], gi[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = t * 16 + acc_row<double>(g, r);
        if (a < N && i < M) {
          const cd v = cmk(gr[t][r], gi[t][r]);
          if (Gw) Gw[(long)a * M + i] = v;
          if (want_e1) e1 += v * Hj[(long)a * M + i];
        }
      }
  }
  if (want_e1) {
    for (int off = 32; off > 0; off >>= 1) {
      e1.re += __shfl_xor(e1.re, off, 64);
      e1.im += __shfl_xor(e1.im, off, 64);
    }
    if (lane == 0) E[(long)w * ldE] = cmk(E0, 0.0) + e1_scale * e1;
  }
}

}  // namespace

void launch_green(hipStream_t s, int nb, int M, int N, const double* PsiT, const double* Phi,
                  long ldw, double* G, double* ov, int ov_mode, double ov_scale, cd ov_add,
                  const double* haj, double E0, double e1_scale, double* E, long ldE) {
  if (nb <= 0 || N <= 0) return;
  AFQ_CHECK(N <= 32, AFQ_EUNSUPPORTED, "mixed density matrix: more than 32 electrons per spin");
  // Phi_w, PsiT and haj go through LDS (batched coalesced loads) when they fit in 96 KB
  const size_t lds = sizeof(cd) * (size_t)M * N * ((E != nullptr && haj != nullptr) ? 3 : 2);
  const bool stage = lds <= 96 * 1024;
  auto go = [&](auto kern) {
    if (stage) ensure_dynamic_lds((const void*)kern, lds);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64), stage ? lds : 0, s, M, N,
                       reinterpret_cast<const cd*>(PsiT), reinterpret_cast<const cd*>(Phi), ldw,
                       reinterpret_cast<cd*>(G), reinterpret_cast<cd*>(ov), ov_mode, ov_scale,
                       ov_add, reinterpret_cast<const cd*>(haj), E0, e1_scale,
                       reinterpret_cast<cd*>(E), ldE);
  };
  if (stage) {
    if (N <= 8) go(k_green<8, true>);
    else if (N <= 16) go(k_green<16, true>);
    else go(k_green<32, true>);
  } else {
    if (N <= 8) go(k_green<8, false>);
    else if (N <= 16) go(k_green<16, false>);
    else go(k_green<32, false>);
  }
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
