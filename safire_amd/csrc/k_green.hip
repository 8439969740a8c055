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

template <int N_>
__device__ inline double ror_d(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_update_dpp((unsigned)u, (unsigned)u, 0x120 + N_, 0xf, 0xf,
                                                  false);
  const unsigned hi = __builtin_amdgcn_update_dpp((unsigned)(u >> 32), (unsigned)(u >> 32),
                                                  0x120 + N_, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int N_>
__device__ inline int ror_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x120 + N_, 0xf, 0xf, false);
}

// maximum over the 16 lanes of each DPP row (every lane ends with its row's maximum; NaN is
// ignored unless a whole row is NaN)
__device__ inline double row_max16(double v) {
  v = fmax(v, ror_d<1>(v));
  v = fmax(v, ror_d<2>(v));
  v = fmax(v, ror_d<4>(v));
  v = fmax(v, ror_d<8>(v));
  return v;
}

__device__ inline int row_min16(int v) {
  v = min(v, ror_i<1>(v));
  v = min(v, ror_i<2>(v));
  v = min(v, ror_i<4>(v));
  v = min(v, ror_i<8>(v));
  return v;
}

__device__ inline int lowest_lane(unsigned long long mask) {
  return __builtin_ctzll(mask);
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
  int rowpos = 0;              // N > 16: the getrf position of this lane's row
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
    // 16 < N <= 32: a row layout that keeps x[] in registers through fully unrolled steps. Lane
    // (r, h) = (lane % 32, lane / 32) holds columns 32h..32h+31 of row r of [T | I] (T in
    // columns 0..31, I in 32..63). Rows are never moved: each row carries its getrf position,
    // the pivot is chosen with izamax's rule over the rows at positions >= j (first maximum of
    // |re|+|im| by position; a NaN at position j is kept, a later NaN is skipped), and positions
    // are swapped as getrf swaps rows, so the eliminations, pivots and signs are getrf's. Each
    // step takes two 16-lane DPP maxima, a few readlanes and one LDS round trip for the pivot
    // row and the multipliers.
    __shared__ cd sRow[2][64];
    __shared__ cd sF[2][32];
    const int r = lane & 31, h = lane >> 5;
#pragma unroll
    for (int e = 0; e < NMAX; ++e) {
      cd v = cmk(0.0, 0.0);
      if (r < N) {
        if (h == 0) {
          if (e < N) v = sT[r * NB + e];
        } else if (e == r) {
          v = cmk(1.0, 0.0);
        }
      }
      x[e] = v;
    }
    int pos = r;
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      if (j < N) {
        const bool cand = h == 0 && r < N && pos >= j;
        const double v = cand ? cabs1(x[j]) : -1.0;
        const int q = lowest_lane(__ballot(h == 0 && pos == j));
        const double vq = readlane_d(v, q);
        int p;
        if (isnan(vq)) {
          p = q;
        } else {
          const double m16 = row_max16(v);
          const double m = fmax(readlane_d(m16, 0), readlane_d(m16, 16));
          const unsigned long long mask = __ballot(cand && v == m);
          if (__builtin_popcountll(mask) == 1) {
            p = lowest_lane(mask);
          } else {
            const int k16 = row_min16(((mask >> lane) & 1ull) ? pos : 0x7fffffff);
            const int km = min(__builtin_amdgcn_readlane(k16, 0), __builtin_amdgcn_readlane(k16, 16));
            p = lowest_lane(__ballot(h == 0 && pos == km));
          }
        }
        p = __builtin_amdgcn_readfirstlane(p);
        const cd piv = cmk(readlane_d(x[j].re, p), readlane_d(x[j].im, p));
        const int pp = __builtin_amdgcn_readlane(pos, p);
        if (lane == j) upiv = (pp == j) ? piv : -piv;   // the sign flips where getrf swaps rows
        if (cis0(piv)) {
          singular = true;   // getrf records the zero pivot and goes on; getri_or_zero zeroes it
        } else {
          if (r == q) pos = pp;
          if (r == p) pos = j;
          cd* row = sRow[j & 1];
          cd* fv = sF[j & 1];
          if (r == p)
#pragma unroll
            for (int e = 0; e < NMAX; ++e) row[32 * h + e] = x[e];
          if (h == 0) fv[r] = x[j];
          wave_sync();
          const cd inv = cdiv(cmk(1.0, 0.0), piv);
          const cd f = fv[r];
          const bool upd = r != p && !cis0(f);
          const cd mlt = f * inv;
          // the pivot row in chunks of 8 columns: each chunk's LDS reads are all in flight before
          // the first use (one wait per chunk rather than one per column)
#pragma unroll
          for (int e0 = 0; e0 < NMAX; e0 += 8) {
            cd pr[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) pr[e] = row[32 * h + e0 + e];
            if (r == p) {
#pragma unroll
              for (int e = 0; e < 8; ++e) x[e0 + e] = pr[e] * inv;
            } else if (upd) {
#pragma unroll
              for (int e = 0; e < 8; ++e) x[e0 + e] -= mlt * pr[e];
            }
          }
        }
      }
    }
    // row r is row pos of [I | T^-1]; lanes (r, 1) publish T^-1 after the zero fill below
    rowpos = pos;
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
  if constexpr (NMAX <= 16) {
    if (lane >= N && lane < 2 * N && !singular) {
      const int a = lane - N;
#pragma unroll
      for (int r = 0; r < NMAX; ++r)
        if (r < N) sT[r * NB + a] = x[r];
    }
  } else {
    if (lane >= 32 && (lane & 31) < N && !singular)
#pragma unroll
      for (int e = 0; e < NMAX; ++e)
        if (e < N) sT[rowpos * NB + e] = x[e];
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


// ---- N <= 16 with the operands staged in LDS: four wavefronts per walker ----
//
// The one-wavefront kernel above runs a serial chain per walker, and a batch of a few hundred
// walkers leaves most SIMDs idle, so its time is that chain's latency. Here a 256-thread
// workgroup shares the walker's work:
//   * all 256 threads stage Phi_w [M][N+1], PsiT [N][M+1] and haj [N][M] in one batch of loads
//     (rows padded so that the 16 rows an MFMA fragment reads fall into distinct banks);
//   * T = PsiT Phi_w is split over k: wavefront v takes the 16-wide k chunks v, v+4, ..., and
//     wavefront 0 sums the four partial tiles in a fixed order;
//   * wavefront 0 inverts T by Gauss-Jordan in a row layout that keeps all 64 lanes busy: lane
//     (r, h) = (lane % 16, lane / 16) holds columns 8h..8h+7 of row r of [T | I] (T in columns
//     0..15, I in 16..31). Rows are never moved: each row carries its getrf position, the pivot
//     is chosen with izamax's rule over the rows at positions >= j (first maximum of |re|+|im|
//     by position; a NaN at position j is kept, a later NaN is skipped), and positions are
//     swapped as getrf swaps rows. The row at position j is then the pivot row of step j, so
//     the eliminations, pivots and signs are getrf's. Each step is one 16-lane DPP maximum,
//     a few readlanes and one LDS round trip for the pivot row and the multipliers;
//   * wavefronts 1-3 compute the 16 x 16 tiles of G (and the E1 partial sums) while wavefront 0
//     takes the logarithms of the pivots.
constexpr int kGWaves = 4;

// e = q d + rem for 0 <= e < 2^24 and 0 < d, from a float estimate corrected by one step
__device__ inline void divmod(int e, int d, float rcp, int& q, int& rem) {
  q = (int)((float)e * rcp);
  rem = e - q * d;
  if (rem < 0) {
    --q;
    rem += d;
  } else if (rem >= d) {
    ++q;
    rem -= d;
  }
}

__global__ __launch_bounds__(64 * kGWaves) void k_green16(
    int M, int N, const cd* __restrict__ PsiT, const cd* __restrict__ Phi, long ldw,
    cd* __restrict__ G, cd* __restrict__ ov, int ov_mode, double ov_scale, cd ov_add,
    const cd* __restrict__ haj, double E0, double e1_scale, cd* __restrict__ E, long ldE) {
  __shared__ cd sPart[kGWaves * 256];      // partial T tiles [wave][a][c]; later T^-1 [b][a]
  __shared__ cd sRow[2][32];               // pivot row of the step (double-buffered)
  __shared__ cd sF[2][16];                 // column j of every row (double-buffered)
  __shared__ cd sE[kGWaves];
  extern __shared__ cd sDyn[];             // Phi_w [M][N+1], PsiT [N][M+1], haj [N][M]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int w = blockIdx.x;
  const bool want_e1 = E != nullptr && haj != nullptr;
  const int ldp = N + 1, ldq = M + 1;
  cd* sPhi = sDyn;
  cd* sPs = sDyn + M * ldp;
  cd* sH = sPs + N * ldq;

  // ---- staging: every load of a batch in flight before the first LDS write ----
  {
    const cd* Pg = Phi + (long)w * ldw;
    const int mn = M * N;
    const int tot = (want_e1 ? 3 : 2) * mn;
    constexpr int B = 12;
    const float rn = 1.0f / (float)N, rm = 1.0f / (float)M;
    for (int e0 = 0; e0 < tot; e0 += 256 * B) {
      double2 v[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int e = min(e0 + 256 * b + tid, tot - 1);
        const cd* src = e < mn ? Pg + e : (e < 2 * mn ? PsiT + (e - mn) : haj + (e - 2 * mn));
        v[b] = *reinterpret_cast<const double2*>(src);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int e = e0 + 256 * b + tid;
        if (e < tot) {
          cd* dst;
          int row, col;
          if (e < mn) {
            divmod(e, N, rn, row, col);
            dst = sPhi + row * ldp + col;
          } else if (e < 2 * mn) {
            divmod(e - mn, M, rm, row, col);
            dst = sPs + row * ldq + col;
          } else {
            dst = sH + (e - 2 * mn);
          }
          *reinterpret_cast<double2*>(dst) = v[b];
        }
      }
    }
  }
  __syncthreads();

  // ---- partial T[a][c] = sum_{i in this wavefront's chunks} PsiT[a][i] Phi[i][c] ----
  {
    d4 tra = d4{0, 0, 0, 0}, trb = d4{0, 0, 0, 0}, tia = d4{0, 0, 0, 0}, tib = d4{0, 0, 0, 0};
    for (int k0 = 16 * wv; k0 < M; k0 += 16 * kGWaves) {
      double ar[4], ai[4], br[4], bi[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = k0 + 4 * g + s;
        const cd x = (r16 < N && k < M) ? sPs[r16 * ldq + k] : cmk(0, 0);
        const cd y = (r16 < N && k < M) ? sPhi[k * ldp + r16] : cmk(0, 0);
        ar[s] = x.re;
        ai[s] = x.im;
        br[s] = y.re;
        bi[s] = y.im;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        tra = mfma16x16x4(ar[s], br[s], tra);
        trb = mfma16x16x4(-ai[s], bi[s], trb);
        tia = mfma16x16x4(ar[s], bi[s], tia);
        tib = mfma16x16x4(ai[s], br[s], tib);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
      sPart[wv * 256 + acc_row<double>(g, r) * 16 + r16] =
          cmk(tra[r] + trb[r], tia[r] + tib[r]);
  }
  __syncthreads();

  cd upiv = cmk(1.0, 0.0);     // lane j < N: +-U_jj
  if (wv == 0) {
    const int r = r16, h = g;
    cd x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int c = 8 * h + q;
      cd v = cmk(0.0, 0.0);
      if (h < 2) {
        if (r < N && c < N) {
          const int e = r * 16 + c;
          v = ((sPart[e] + sPart[256 + e]) + sPart[512 + e]) + sPart[768 + e];
        }
      } else if (r < N && c - 16 == r) {
        v = cmk(1.0, 0.0);
      }
      x[q] = v;
    }
    int pos = r;                 // getrf position of row r
    bool singular = false;
    // steps j >= N do nothing: a uniform branch rather than a break, so that the loop unrolls
    // and x[] stays in registers
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if (j < N) {
        const int hj = j >> 3, qj = j & 7;
        const bool cand = h == hj && r < N && pos >= j;
        const double v = cand ? cabs1(x[qj]) : -1.0;
        // the row now at position j, and izamax's choice
        const int q = lowest_lane(__ballot(h == 0 && pos == j));
        const double vq = readlane_d(v, q + 16 * hj);
        int p;
        if (isnan(vq)) {
          p = q;
        } else {
          const double m = row_max16(v);
          const unsigned long long mask = __ballot(cand && v == m);
          if (__builtin_popcountll(mask) == 1) {
            p = lowest_lane(mask) - 16 * hj;
          } else {
            const int kmin = row_min16(((mask >> lane) & 1ull) ? pos : 0x7fffffff);
            const int km = __builtin_amdgcn_readlane(kmin, 16 * hj);
            p = lowest_lane(__ballot(h == 0 && pos == km));
          }
        }
        p = __builtin_amdgcn_readfirstlane(p);
        const int lp = p + 16 * hj;
        const cd piv = cmk(readlane_d(x[qj].re, lp), readlane_d(x[qj].im, lp));
        const int pp = __builtin_amdgcn_readlane(pos, lp);
        if (lane == j) upiv = (pp == j) ? piv : -piv;   // the sign flips where getrf swaps rows
        if (cis0(piv)) {
          singular = true;   // getrf records the zero pivot and goes on; getri_or_zero zeroes it
        } else {
          if (r == q) pos = pp;
          if (r == p) pos = j;
          cd* row = sRow[j & 1];
          cd* fv = sF[j & 1];
          if (r == p)
#pragma unroll
            for (int e = 0; e < 8; ++e) row[8 * h + e] = x[e];
          if (h == hj) fv[r] = x[qj];
          wave_sync();
          const cd inv = cdiv(cmk(1.0, 0.0), piv);
          cd pr[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) pr[e] = row[8 * h + e];
          const cd f = fv[r];
          if (r == p) {
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = pr[e] * inv;
          } else if (!cis0(f)) {
            const cd mlt = f * inv;
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] -= mlt * pr[e];
          }
        }
      }
    }
    // row r is row pos of [I | T^-1]: T^-1 to LDS as [b][a]
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      if (h >= 2) {
        const int c = 8 * (h - 2) + e;
        sPart[pos * 16 + c] = (!singular && r < N && c < N) ? x[e] : cmk(0.0, 0.0);
      }
    }
  }
  __syncthreads();

  if (wv == 0) {
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
  }
  if (G == nullptr && E == nullptr) return;

  // ---- G[a][i] = sum_b Tinv[b][a] Phi[i][b], one 16-column tile of i per step ----
  cd e1 = cmk(0.0, 0.0);
  cd* Gw = G ? G + (long)w * N * M : nullptr;
  if (wv > 0) {
    for (int i0 = 16 * (wv - 1); i0 < M; i0 += 16 * (kGWaves - 1)) {
      d4 gra = d4{0, 0, 0, 0}, grb = d4{0, 0, 0, 0}, gia = d4{0, 0, 0, 0}, gib = d4{0, 0, 0, 0};
      const int i = i0 + r16;
      double tr_[4], ti_[4], pr_[4], pi_[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int b = 4 * g + s;
        const cd y = (i < M && b < N) ? sPhi[i * ldp + b] : cmk(0, 0);
        pr_[s] = y.re;
        pi_[s] = y.im;
        const cd z = sPart[b * 16 + r16];            // Tinv^T[a][b] = Tinv[b][a]
        tr_[s] = z.re;
        ti_[s] = z.im;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        gra = mfma16x16x4(tr_[s], pr_[s], gra);
        grb = mfma16x16x4(-ti_[s], pi_[s], grb);
        gia = mfma16x16x4(tr_[s], pi_[s], gia);
        gib = mfma16x16x4(ti_[s], pr_[s], gib);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = acc_row<double>(g, r);
        if (a < N && i < M) {
          const cd v = cmk(gra[r] + grb[r], gia[r] + gib[r]);
          if (Gw) Gw[(long)a * M + i] = v;
          if (want_e1) e1 += v * sH[a * M + i];
        }
      }
    }
  }
  if (want_e1) {
    for (int off = 32; off > 0; off >>= 1) {
      e1.re += __shfl_xor(e1.re, off, 64);
      e1.im += __shfl_xor(e1.im, off, 64);
    }
    if (lane == 0) sE[wv] = e1;
    __syncthreads();
    if (tid == 0) E[(long)w * ldE] = cmk(E0, 0.0) + e1_scale * ((sE[1] + sE[2]) + sE[3]);
  }
}

}  // namespace

void launch_green(hipStream_t s, int nb, int M, int N, const double* PsiT, const double* Phi,
                  long ldw, double* G, double* ov, int ov_mode, double ov_scale, cd ov_add,
                  const double* haj, double E0, double e1_scale, double* E, long ldE) {
  if (nb <= 0 || N <= 0) return;
  AFQ_CHECK(N <= 32, AFQ_EUNSUPPORTED, "mixed density matrix: more than 32 electrons per spin");
  const bool e1 = E != nullptr && haj != nullptr;
  // N <= 16: four wavefronts per walker when Phi_w, PsiT and haj (rows padded) fit in 96 KB
  const size_t lds16 = sizeof(cd) * ((size_t)M * (N + 1) + (size_t)N * (M + 1) +
                                     (e1 ? (size_t)M * N : 0));
  if (N <= 16 && lds16 <= 96 * 1024) {
    ensure_dynamic_lds((const void*)k_green16, lds16);
    hipLaunchKernelGGL(k_green16, dim3(nb), dim3(64 * kGWaves), lds16, s, M, N,
                       reinterpret_cast<const cd*>(PsiT), reinterpret_cast<const cd*>(Phi), ldw,
                       reinterpret_cast<cd*>(G), reinterpret_cast<cd*>(ov), ov_mode, ov_scale,
                       ov_add, reinterpret_cast<const cd*>(haj), E0, e1_scale,
                       reinterpret_cast<cd*>(E), ldE);
    AFQ_HIP(hipGetLastError());
    return;
  }
  // Phi_w, PsiT and haj go through LDS (batched coalesced loads) when they fit in 96 KB
  const size_t lds = sizeof(cd) * (size_t)M * N * (e1 ? 3 : 2);
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
    go(k_green<32, true>);
  } else {
    if (N <= 8) go(k_green<8, false>);
    else if (N <= 16) go(k_green<16, false>);
    else go(k_green<32, false>);
  }
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
