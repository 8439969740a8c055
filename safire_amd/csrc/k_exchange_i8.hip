// Exchange and Coulomb energy (Real3IndexFactorization::energy_impl, Real3IndexFactorization.hpp:
// 476-559) with the contraction T_wn = G_w Lnak_n^T on int8 MFMA (v_mfma_i32_16x16x64_i8),
// split after Ozaki: every row of G and of each Lnak_n is scaled by a power of two u and
// written as a sum of int8 slices, x = u (q0 + q1/254 + q2/254^2 + ...) with every |qs| <= 127
// (q0 = rint(x/u), |x/u| < 127; each residual lies in [-1/2, 1/2], times 254 in [-127, 127]).
// A slice product accumulates in int32 without rounding, so the error left is the truncation
// of the slice series, relative to the row maximum:
//
// * 5 slices of G and 5 of Lnak (remainder 254^-4 / 2 of the unit, ~2^-39 of the row maximum).
//   The truncation of G is the same for every Cholesky vector and adds up coherently over the
//   K vectors; the per-row scale costs ~4 bits against FP32's per-element exponent.
// * All 15 slice products (s, t) with s + t <= 4 are kept; dropping any product of order 4
//   raises the energy error to ~0.1 of the FP32 form's. Products of one order share the scale
//   254^-(s+t) and accumulate in the same int32 registers (|sum| < 2^25 for M <= 128);
//   T = uG_a uL_b sum_o 254^-o acc_o is assembled in FP64.
//
// On the synthetic Hamiltonians the energy error is ~1e-3 of the FP32 form's. The MFMA work is
// 60 int8 16x16x64 products per walker, vector and 64-wide k chunk.
//
// Structure: a workgroup holds WPB walkers (one per wavefront) and a chunk of Cholesky vectors.
// Each wavefront slices its walker's G once into A fragments kept in registers (Re, Im and
// -Im); the Lnak slices (precomputed at set-up) of kStage vectors are staged in LDS per group;
// each N x N block of T goes through per-wavefront LDS for tr(T^2) and tr(T) in FP64.
// N <= 16 (one 16x16 tile per block), M <= 128.
#include "afq_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace afq {

namespace {

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int kSG = 5;      // slices of G
constexpr int kSL = 5;      // slices of Lnak
constexpr int kOrders = 5;  // accumulator orders 0..4
// walkers (wavefronts) per workgroup: two 4-wavefront workgroups per CU run out of step, so
// one's FP64 epilogue overlaps the other's MFMAs (one 8-wavefront workgroup keeps all of a
// CU's wavefronts in step at its group barrier)
constexpr int kWPB = 4;
constexpr int kWgPerCu = 2;   // resident workgroups per CU (LDS and registers allow two)
constexpr int kMaxKC = 2;   // 64-wide k chunks: M <= 128
constexpr double kBase = 254.0;   // ratio of consecutive slice units

// kept slice products: all of order s + t <= 4
__host__ __device__ constexpr bool kept(int s, int t) { return s + t < kOrders; }
// vectors staged per group: LDS holds 2 buffers x kStage x 10 slice rows of 64 KC (+16) bytes
// x 16 rows, plus the per-wavefront T blocks
__host__ __device__ constexpr int stage_of(int KC) { return KC == 1 ? 2 : 1; }

// power-of-two unit of a row whose largest |component| is mx: |x / u| < 127 for every x of the
// row. Rows below 2^-960 count as zero (their products are far below any energy resolution);
// a non-finite row gets NaN so that its energies come out NaN as in FP64.
__device__ inline double row_unit(double mx, bool& zero) {
  zero = false;
  if (!(mx <= 1.0e300)) return __builtin_nan("");
  if (mx < 0x1p-960) {
    zero = true;
    return 1.0;
  }
  int e;
  frexp(mx * (128.0 / 127.0), &e);
  return ldexp(1.0, e - 7);
}

// the first S slices of r = x / u (|r| < 127; x * inv_u is exact, inv_u being a power of two)
template <int S>
__device__ inline void slice(double r, int (&q)[S]) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const double qs = rint(r);
    q[s] = (int)qs;
    r = (r - qs) * kBase;
  }
}

__device__ inline unsigned pack4(int a, int b, int c, int d) {
  return (unsigned)(a & 0xff) | ((unsigned)(b & 0xff) << 8) | ((unsigned)(c & 0xff) << 16) |
         ((unsigned)(d & 0xff) << 24);
}

__device__ inline i4 mfma_i8(i4 a, i4 b, i4 c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// v plus the value of the lane `shift` below in the same row of 16 lanes (0 below the row)
template <int SHIFT>
__device__ inline double add_row_shr(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)u, 0x110 + SHIFT, 0xf, 0xf, true);
  const unsigned hi =
      __builtin_amdgcn_update_dpp(0u, (unsigned)(u >> 32), 0x110 + SHIFT, 0xf, 0xf, true);
  return v + __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// lane 15 of each row of 16 lanes receives the sum of the row (DPP row shifts, no LDS)
__device__ inline double row_sum16(double v) {
  v = add_row_shr<1>(v);
  v = add_row_shr<2>(v);
  v = add_row_shr<4>(v);
  return add_row_shr<8>(v);
}

__device__ inline double readlane_d(double v, int lane) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// orders LDS traffic of one wavefront: its own writes are visible to its own later reads
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// workgroup barrier after this wavefront's LDS reads and writes have completed; global loads and
// copies into LDS still in flight are not waited for (a fence would wait for them: the copies
// write LDS). Every access to the copies' destination waits for them explicitly (vmcnt) before
// the next full barrier. The memory clobber keeps the compiler from moving LDS accesses across.
__device__ inline void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int KC, int NV>
__global__ __launch_bounds__(64 * kWPB, kWgPerCu * kWPB / 4) void k_exchange_i8(
    int W, int M, int N, int K, int Mp, int kChunk, const double* __restrict__ G,
    const int8_t* __restrict__ Lq, const double* __restrict__ Lu, double* __restrict__ Kl,
    double* __restrict__ exx_part, int nchunk) {
  constexpr int ROWB = 64 * KC + 16;   // staged row stride (bytes): the row plus 16 of padding
  constexpr int kStage = stage_of(KC);
  constexpr int kGroupVec = kStage * kSL * 2 * 16 * (4 * KC);   // 16-byte vectors per group
  constexpr int kPre = (kGroupVec + 64 * kWPB - 1) / (64 * kWPB);  // ... per thread
  // two staging buffers: the next group is loaded while the current one is computed on
  __shared__ __attribute__((aligned(16))) int8_t sBuf[2][kStage * kSL * 2 * 16 * ROWB];
  __shared__ double sUnit[2][kStage][16];           // column units of the staged vectors
  // per-wavefront T block [a][b] (re, im), rows padded to 17 entries: the transposed read
  // (16 lanes down a column) then falls into distinct banks
  __shared__ __attribute__((aligned(16))) double sT[kWPB][16 * 17 * 2];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int w_raw = blockIdx.x * kWPB + wave;
  const bool active = w_raw < W;
  const int w = active ? w_raw : W - 1;
  const int c = blockIdx.y;
  const int nbeg = c * kChunk;
  const int nend = min(K, nbeg + kChunk);
  const long KNMp = (long)K * N * Mp;

  // ---- slices of G_w: row a = r16, k = 64 kc + 16 g + j (j < 16)
  i4 ar[KC][kSG], ai[KC][kSG], nai[KC][kSG];
  double uG;
  {
    const double* Gw = G + 2 * ((long)w * N + r16) * M;
    double mx = 0.0;
    if (r16 < N) {
      for (int kc = 0; kc < KC; ++kc)
        for (int j = 0; j < 16; ++j) {
          const int k = 64 * kc + 16 * g + j;
          if (k < M) {
            const double xr = fabs(Gw[2 * k]), xi = fabs(Gw[2 * k + 1]);
            // a NaN or Inf anywhere in the row must survive the max (fmax drops NaN)
            mx = (xr <= 1.0e300 && xi <= 1.0e300) ? fmax(mx, fmax(xr, xi)) : INFINITY;
          }
        }
    }
    mx = fmax(mx, __shfl_xor(mx, 16, 64));
    mx = fmax(mx, __shfl_xor(mx, 32, 64));
    bool zero;
    uG = row_unit(mx, zero);
    const bool finite = uG == uG;
    const double inv = (finite && !zero) ? 1.0 / uG : 0.0;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
      for (int d = 0; d < 4; ++d) {   // bytes 4d..4d+3 of each fragment
        int qr[4][kSG], qi[4][kSG];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = 64 * kc + 16 * g + 4 * d + j;
          double xr = 0.0, xi = 0.0;
          if (r16 < N && k < M && inv != 0.0) {
            xr = Gw[2 * k] * inv;
            xi = Gw[2 * k + 1] * inv;
          }
          slice<kSG>(xr, qr[j]);
          slice<kSG>(xi, qi[j]);
        }
#pragma unroll
        for (int s = 0; s < kSG; ++s) {
          ar[kc][s][d] = (int)pack4(qr[0][s], qr[1][s], qr[2][s], qr[3][s]);
          ai[kc][s][d] = (int)pack4(qi[0][s], qi[1][s], qi[2][s], qi[3][s]);
          nai[kc][s][d] = (int)pack4(-qi[0][s], -qi[1][s], -qi[2][s], -qi[3][s]);
        }
      }
    }
  }
  // units of the accumulator rows 4g + r (row a's unit is held by lane a)
  double uRow[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) uRow[r] = __shfl(uG, 4 * g + r, 64);

  double exx_re = 0.0, exx_im = 0.0;
  double* T = sT[wave];
  const int nvec = 4 * KC;          // 16-byte vectors per staged row

  // group loads: slice rows of vectors nv0 .. nv0 + kStage - 1 into registers (zero outside).
  // Item e = i * 64 kWPB + tid is (row b, 16-byte column kv) = pos, which is fixed per thread
  // (16 nvec divides 64 kWPB), and (vector v, slice-part tp) = combo, which is fixed per
  // wavefront and slot, so only the thread's part of the address lives in vector registers.
  static_assert((64 * kWPB) % (16 * 4 * KC) == 0, "staging positions must tile the workgroup");
  const int pos = threadIdx.x % (16 * nvec);
  const int pb = pos / nvec, pkv = pos % nvec;
  const bool pos_ok = pb < N && 16 * pkv < Mp;
  const long gpos = (long)pb * Mp + 16 * pkv;
  const int lpos = pb * ROWB + 16 * pkv;
  // per slot i: combo = v * 2 kSL + tp (wave-uniform) and the uniform part of its address,
  // tp (K N Mp) + v (N Mp); a group only adds (nv0 N Mp) and the thread's gpos. KC = 1 keeps
  // them in registers across groups; KC = 2 has no registers to spare and recomputes them.
  const long NMp = (long)N * Mp;
  int combo_k[kPre], vsl_k[kPre], ubase_k[kPre];   // offsets fit in 32 bits (checked at launch)
  auto slot = [&](int i, int& combo, int& v, int& ubase) {
    if constexpr (KC == 1) {
      combo = combo_k[i];
      v = vsl_k[i];
      ubase = ubase_k[i];
    } else {
      combo = __builtin_amdgcn_readfirstlane((i * 64 * kWPB + (int)threadIdx.x) / (16 * nvec));
      v = combo / (kSL * 2);
      ubase = (int)((combo % (kSL * 2)) * KNMp) + v * (int)NMp;
    }
  };
  if constexpr (KC == 1) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      combo_k[i] =
          __builtin_amdgcn_readfirstlane((i * 64 * kWPB + (int)threadIdx.x) / (16 * nvec));
      vsl_k[i] = combo_k[i] / (kSL * 2);
      ubase_k[i] = (int)((combo_k[i] % (kSL * 2)) * KNMp) + vsl_k[i] * (int)NMp;
    }
  }
  uint4 pre[kPre];
  double preU = 0.0;
  auto load_group = [&](int nv0) {
    const int8_t* Lg = Lq + (long)nv0 * NMp + gpos;
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      int combo, v, ubase;
      slot(i, combo, v, ubase);
      pre[i] = uint4{0u, 0u, 0u, 0u};
      if (combo < kStage * kSL * 2 && nv0 + v < nend && pos_ok) {
        if constexpr (KC == 1)
          pre[i] = *reinterpret_cast<const uint4*>(Lg + ubase);
        else
          pre[i] = *reinterpret_cast<const uint4*>(Lq + (combo % (kSL * 2)) * KNMp +
                                                    (long)(nv0 + v) * NMp + gpos);
      }
    }
    if (threadIdx.x < kStage * 16) {
      const int v = threadIdx.x / 16, b = threadIdx.x % 16, n = nv0 + v;
      preU = (n < nend && b < N) ? Lu[(long)n * N + b] : 0.0;
    }
  };
  auto store_group = [&](int buf) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      int combo, v, ubase;
      slot(i, combo, v, ubase);
      if (combo < kStage * kSL * 2)   // rows of slice-part combo
        *reinterpret_cast<uint4*>(sBuf[buf] + combo * 16 * ROWB + lpos) = pre[i];
    }
    if (threadIdx.x < kStage * 16) sUnit[buf][threadIdx.x / 16][threadIdx.x % 16] = preU;
  };

  if (nbeg < nend) {
    load_group(nbeg);
    store_group(0);
  }
  __syncthreads();
  for (int nv0 = nbeg, buf = 0; nv0 < nend; nv0 += kStage, buf ^= 1) {
    const bool more = nv0 + kStage < nend;
    if (more) load_group(nv0 + kStage);   // in flight during this group's MFMAs
    const int8_t* sB = sBuf[buf];
    const double (*sU)[16] = sUnit[buf];

#pragma unroll 1
    for (int h = 0; h < kStage; h += NV) {
      i4 accR[NV][kOrders], accI[NV][kOrders];
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int o = 0; o < kOrders; ++o) {
          accR[v][o] = i4{0, 0, 0, 0};
          accI[v][o] = i4{0, 0, 0, 0};
        }
#pragma unroll
      for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int t = 0; t < kSL; ++t) {
            const int8_t* base = sB + (((h + v) * kSL + t) * 2 * 16 + r16) * ROWB + 64 * kc + 16 * g;
            const i4 br = *reinterpret_cast<const i4*>(base);
            const i4 bi = *reinterpret_cast<const i4*>(base + 16 * ROWB);
#pragma unroll
            for (int s = 0; s < kSG; ++s)
              if (kept(s, t)) {
                accR[v][s + t] = mfma_i8(ar[kc][s], br, accR[v][s + t]);
                accI[v][s + t] = mfma_i8(ar[kc][s], bi, accI[v][s + t]);
              }
#pragma unroll
            for (int s = 0; s < kSG; ++s)
              if (kept(s, t)) {
                accR[v][s + t] = mfma_i8(nai[kc][s], bi, accR[v][s + t]);
                accI[v][s + t] = mfma_i8(ai[kc][s], br, accI[v][s + t]);
              }
          }

      // epilogue in FP64: T block -> LDS, tr(T^2) lane partials, tr(T) -> Kl
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int n = nv0 + h + v;
        const double uL = sU[h + v][r16];
        double tr[4], ti[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (KC == 1) {
            // M <= 64: |acc_o| <= (o + 1) 2 64 127^2, so 254 acc0 + acc1 and 254 acc2 + acc3
            // stay below 2^31 and pair up exactly in int32 (three conversions, not five)
            // (24-bit multiplies: |acc0|, |acc2| < 2^23)
            const int pr01 = __mul24(254, accR[v][0][r]) + accR[v][1][r];
            const int pi01 = __mul24(254, accI[v][0][r]) + accI[v][1][r];
            const int pr23 = __mul24(254, accR[v][2][r]) + accR[v][3][r];
            const int pi23 = __mul24(254, accI[v][2][r]) + accI[v][3][r];
            double xr = fma((double)accR[v][4][r], 1.0 / kBase, (double)pr23);
            double xi = fma((double)accI[v][4][r], 1.0 / kBase, (double)pi23);
            xr = fma(xr, 1.0 / (kBase * kBase), (double)pr01);
            xi = fma(xi, 1.0 / (kBase * kBase), (double)pi01);
            const double u = uRow[r] * uL * (1.0 / kBase);   // pr01 carries one factor 254
            tr[r] = xr * u;
            ti[r] = xi * u;
          } else {
            double xr = (double)accR[v][kOrders - 1][r], xi = (double)accI[v][kOrders - 1][r];
#pragma unroll
            for (int o = kOrders - 2; o >= 0; --o) {
              xr = fma(xr, 1.0 / kBase, (double)accR[v][o][r]);
              xi = fma(xi, 1.0 / kBase, (double)accI[v][o][r]);
            }
            const double u = uRow[r] * uL;
            tr[r] = xr * u;
            ti[r] = xi * u;
          }
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rw = acc_row<float>(g, r);
          *reinterpret_cast<double2*>(T + 2 * (rw * 17 + r16)) = double2{tr[r], ti[r]};
        }
        wave_lds_sync();
        if (n < nend) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rw = acc_row<float>(g, r);
            const double2 y = *reinterpret_cast<const double2*>(T + 2 * (r16 * 17 + rw));
            const double yr = y.x, yi = y.y;
            exx_re += tr[r] * yr - ti[r] * yi;
            exx_im += tr[r] * yi + ti[r] * yr;
          }
          // tr(T): the N diagonal entries in lanes 0..15, summed into lane 15 by DPP row shifts
          double dr = 0.0, di = 0.0;
          if (lane < N) {
            const double2 d = *reinterpret_cast<const double2*>(T + 2 * (lane * 17 + lane));
            dr = d.x;
            di = d.y;
          }
          dr = row_sum16(dr);
          di = row_sum16(di);
          if (lane == 15 && active) {
            Kl[2 * ((long)w * K + n)] = dr;
            Kl[2 * ((long)w * K + n) + 1] = di;
          }
        }
      }
    }
    // the other buffer was last read in the previous group, before the barrier that ended it
    if (more) store_group(buf ^ 1);
    __syncthreads();
  }

  for (int off = 32; off > 0; off >>= 1) {
    exx_re += __shfl_xor(exx_re, off, 64);
    exx_im += __shfl_xor(exx_im, off, 64);
  }
  if (lane == 0 && active) {
    exx_part[2 * ((long)w * nchunk + c)] = exx_re;
    exx_part[2 * ((long)w * nchunk + c) + 1] = exx_im;
  }
}

// N in 17..32: T_wn is 2 x 2 tiles of 16 x 16. An 8-wavefront workgroup holds 4 walkers; the
// two wavefronts of a walker own its row tiles ta = 0, 1 (G rows 16 ta .. 16 ta + 15 sliced into
// registers) and compute both column tiles of their rows.
//
// Slices: 6 of G and 6 of each Lnak row (base 254), and per Cholesky vector a slice-product order
// o_n in 2..5 (all products s + t <= o_n kept; 21, 15, 10 or 6 products) fixed at set-up from
// the vector's size r_n = max|L_n| / max_m |L_m|: o_n = ceil(5 + 2 log_254 r_n). The error a
// vector adds to EXX scales as r_n^2 254^-(o_n + 1) (both its T block and the truncation of its
// product series are proportional to r_n), so every vector then contributes at most 254^-6 of
// the largest one's scale: on Cholesky vectors whose sizes decay, the large vectors take all 21
// products and the small ones few, fewer products on average than a uniform order 4 (15) for
// tens of times less error. The real part keeps Re G Re L and Im G Im L in separate int32
// accumulators (subtracted exactly before the FP64 assembly), so no negated slices are needed.
//
// tr(T^2) = tr(T00^2) + tr(T11^2) + 2 tr(T01 T10): wavefront ta = 0 takes the first and the cross
// term (it holds T00 and T01 in registers and reads T00^T and T10 from LDS), wavefront 1 the
// second (T11^T from LDS); so only T00, T10 and T11 go to LDS, and wavefront 0 of the walker
// takes tr(T) from their diagonals. The two row tiles write separate tr(T^2) partials (slots
// 2c and 2c + 1 of the walker).
//
// Staging: the int8 slices of one vector (12 slice-parts x 32 rows x 64 KC bytes, 24 or 48 KB)
// are copied from global memory straight into LDS (global_load_lds_dwordx4, no registers),
// double-buffered in two LDS objects with the vector loop unrolled by two: every access then
// names its buffer statically, so the compiler does not wait for the next vector's copies before
// reading the current one. Rows are not padded; the 16-byte chunk kv of row b sits at position
// kv ^ f(b) of its row (XOR swizzle), which spreads the 16 rows an MFMA fragment reads over
// distinct banks. Rows b >= N and chunks past the end of a global row copy clamped (in-bounds)
// data: their column unit is 0 and the matching G entries are 0, so they contribute nothing.
// LDS: 2 x 48 KB + 4 x 12 KB at M = 128 (one workgroup, two wavefronts per SIMD).
constexpr int kT2Walkers = 4;
constexpr int kS6 = 6;          // slices of G and of Lnak (N > 16)
constexpr int kO6 = 6;          // accumulator orders 0..5

template <int KC>
__device__ inline int t2_swz(int b) {
  return KC == 2 ? ((b >> 1) & 7) : ((b >> 2) & 3);
}

template <int KC>
__global__ __launch_bounds__(128 * kT2Walkers, 2) void k_exchange_i8_t2(
    int W, int M, int N, int K, int Mp, int kChunk, const double* __restrict__ G,
    const int8_t* __restrict__ Lq, const double* __restrict__ Lu, const int* __restrict__ ord,
    double* __restrict__ Kl, double* __restrict__ exx_part, int nslots) {
  constexpr int kRows = 32;
  constexpr int ROWB = 64 * KC;                  // staged row (bytes), unpadded
  constexpr int kPlane = kRows * ROWB;           // one slice-part of one vector
  constexpr int kVecBytes = kS6 * 2 * kPlane;
  constexpr int kWindows = kVecBytes / 1024;     // 1 KB copied per wavefront instruction
  constexpr int kWaves = 2 * kT2Walkers;
  constexpr int kTile = 16 * 16 * 2;             // doubles of one 16 x 16 complex tile
  static_assert(kVecBytes % 1024 == 0, "a vector's slices are whole 1 KB windows");
  __shared__ __attribute__((aligned(16))) int8_t sB0[kVecBytes];
  __shared__ __attribute__((aligned(16))) int8_t sB1[kVecBytes];
  __shared__ double sU0[kRows];
  __shared__ double sU1[kRows];
  // per walker the tiles T00, T10, T11 [x][y] (re, im); entry (x, y) sits in column y ^ x of row
  // x, so that 16 lanes reading down a column hit 16 different banks
  __shared__ double sT[kT2Walkers][3 * kTile];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int wl = wave >> 1, ta = wave & 1;
  const int w_raw = blockIdx.x * kT2Walkers + wl;
  const bool active = w_raw < W;
  const int w = active ? w_raw : W - 1;
  const int c = blockIdx.y;
  const int nbeg = c * kChunk;
  const int nend = min(K, nbeg + kChunk);
  const long KNMp = (long)K * N * Mp;
  const int arow = ta * 16 + r16;                   // this lane's row of G

  i4 ar[KC][kS6], ai[KC][kS6];
  double uG;
  {
    const double* Gw = G + 2 * ((long)w * N + arow) * M;
    double mx = 0.0;
    if (arow < N) {
      for (int kc = 0; kc < KC; ++kc)
        for (int j = 0; j < 16; ++j) {
          const int k = 64 * kc + 16 * g + j;
          if (k < M) {
            const double xr = fabs(Gw[2 * k]), xi = fabs(Gw[2 * k + 1]);
            mx = (xr <= 1.0e300 && xi <= 1.0e300) ? fmax(mx, fmax(xr, xi)) : INFINITY;
          }
        }
    }
    mx = fmax(mx, __shfl_xor(mx, 16, 64));
    mx = fmax(mx, __shfl_xor(mx, 32, 64));
    bool zero;
    uG = row_unit(mx, zero);
    const bool finite = uG == uG;
    const double inv = (finite && !zero) ? 1.0 / uG : 0.0;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        int qr[4][kS6], qi[4][kS6];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = 64 * kc + 16 * g + 4 * d + j;
          double xr = 0.0, xi = 0.0;
          if (arow < N && k < M && inv != 0.0) {
            xr = Gw[2 * k] * inv;
            xi = Gw[2 * k + 1] * inv;
          }
          slice<kS6>(xr, qr[j]);
          slice<kS6>(xi, qi[j]);
        }
#pragma unroll
        for (int s = 0; s < kS6; ++s) {
          ar[kc][s][d] = (int)pack4(qr[0][s], qr[1][s], qr[2][s], qr[3][s]);
          ai[kc][s][d] = (int)pack4(qi[0][s], qi[1][s], qi[2][s], qi[3][s]);
        }
      }
    }
  }
  double uRow[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) uRow[r] = __shfl(uG, 4 * g + r, 64);

  double exx_re = 0.0, exx_im = 0.0;
  double* T = sT[wl];
  // LDS tile of (row tile x, column tile y): T00 -> 0, T10 -> 1, T11 -> 2 (T01 stays in registers)
  auto tslot = [](int x, int y) { return x == 0 ? 0 : 1 + y; };
  auto tpos = [](int x, int y) { return 2 * (x * 16 + (y ^ x)); };   // within a tile

  // copy vector n's slices into dst: window win of the vector is LDS bytes
  // [1024 win, 1024 win + 1024), lane l writes bytes 16 l .. 16 l + 15 of it
  auto stage = [&](int n, int8_t* dst) {
    for (int win = wave; win < kWindows; win += kWaves) {
      const int o = win * 1024 + 16 * lane;
      const int tp = o / kPlane;
      const int b = (o % kPlane) / ROWB;
      const int kv = ((o % ROWB) >> 4) ^ t2_swz<KC>(b);
      const int bg = min(b, N - 1);
      const int kg = min(16 * kv, Mp - 16);
      const int8_t* src = Lq + tp * KNMp + ((long)n * N + bg) * Mp + kg;
      __builtin_amdgcn_global_load_lds(
          (const void __attribute__((address_space(1)))*)src,
          (void __attribute__((address_space(3)))*)(dst + win * 1024), 16, 0, 0);
    }
  };
  double preU = 0.0;
  auto load_units = [&](int n) {
    if (threadIdx.x < kRows) preU = (int)threadIdx.x < N ? Lu[(long)n * N + threadIdx.x] : 0.0;
  };

  if (nbeg < nend) {
    load_units(nbeg);
    stage(nbeg, sB0);
  }
  // one vector: its slices are in sBb (or in flight to it) and its column units in preU; they
  // go to uB here, and the next vector's slices to sBn
  auto vector = [&](int n, const int8_t* sBb, double* uB, int8_t* sBn) {
    // vector n is in sBb (this thread's copies done, then everyone's); every wavefront is past
    // vector n - 1, so sBn, uB and the T tiles are free again
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    if (threadIdx.x < kRows) uB[threadIdx.x] = preU;
    __syncthreads();
    const int on = __builtin_amdgcn_readfirstlane(ord[n]);   // this vector's slice order
    if (n + 1 < nend) {
      load_units(n + 1);   // before the copies: waiting for it leaves them in flight
      stage(n + 1, sBn);
    }
    double kre[4], kim[4];   // T01 (rows of wavefront 0, column tile 1) stays in registers
#pragma unroll
    for (int tb = 0; tb < 2; ++tb) {
      i4 acR[kO6], acQ[kO6], acI[kO6];   // Re G Re L, Im G Im L, Re G Im L + Im G Re L by order
#pragma unroll
      for (int o = 0; o < kO6; ++o) {
        acR[o] = i4{0, 0, 0, 0};
        acQ[o] = i4{0, 0, 0, 0};
        acI[o] = i4{0, 0, 0, 0};
      }
      const int row = tb * 16 + r16;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
        const int chunk = (((4 * kc + g) ^ t2_swz<KC>(row)) << 4);
#pragma unroll
        for (int t = 0; t < kS6; ++t) {
          if (t > on) break;   // uniform over the workgroup
          const int8_t* base = sBb + ((t * 2) * kRows + row) * ROWB + chunk;
          const i4 br = *reinterpret_cast<const i4*>(base);
          const i4 bi = *reinterpret_cast<const i4*>(base + kPlane);
#pragma unroll
          for (int s = 0; s < kS6; ++s)
            if (s + t < kO6 && s + t <= on) {
              acR[s + t] = mfma_i8(ar[kc][s], br, acR[s + t]);
              acI[s + t] = mfma_i8(ar[kc][s], bi, acI[s + t]);
              acQ[s + t] = mfma_i8(ai[kc][s], bi, acQ[s + t]);
              acI[s + t] = mfma_i8(ai[kc][s], br, acI[s + t]);
            }
        }
      }
      const double uL = uB[row];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double xr = (double)(acR[kO6 - 1][r] - acQ[kO6 - 1][r]);
        double xi = (double)acI[kO6 - 1][r];
#pragma unroll
        for (int o = kO6 - 2; o >= 0; --o) {
          xr = fma(xr, 1.0 / kBase, (double)(acR[o][r] - acQ[o][r]));
          xi = fma(xi, 1.0 / kBase, (double)acI[o][r]);
        }
        const double u = uRow[r] * uL;
        if (ta == 0 && tb == 1) {
          kre[r] = xr * u;
          kim[r] = xi * u;
        } else {
          double* tile = T + tslot(ta, tb) * kTile;
          const int x = 4 * g + r, y = r16;
          tile[tpos(x, y)] = xr * u;
          tile[tpos(x, y) + 1] = xi * u;
        }
      }
    }
    lds_barrier();   // T00, T10, T11 of every walker are in LDS
    // ta = 0: sum_{a,b in tile 0} T_ab T_ba + 2 sum_{a in 0, b in 1} T_ab T_ba;
    // ta = 1: sum_{a,b in tile 1} T_ab T_ba. For (a, b) = (4 g + r, r16) within the tiles, T_ab
    // is entry (4 g + r, r16) of its tile and T_ba entry (r16, 4 g + r) of the tile (row tile of
    // b, column tile of a).
    {
      const double* diag = T + (ta == 0 ? 0 : 2) * kTile;   // T00 or T11
      const double* cross = T + 1 * kTile;                 // T10
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = 4 * g + r, y = r16;
        const double xr = diag[tpos(x, y)], xi = diag[tpos(x, y) + 1];
        const double yr = diag[tpos(y, x)], yi = diag[tpos(y, x) + 1];
        exx_re += xr * yr - xi * yi;
        exx_im += xr * yi + xi * yr;
        if (ta == 0) {
          const double zr = cross[tpos(y, x)], zi = cross[tpos(y, x) + 1];
          exx_re += 2.0 * (kre[r] * zr - kim[r] * zi);
          exx_im += 2.0 * (kre[r] * zi + kim[r] * zr);
        }
      }
    }
    if (ta == 0) {
      // tr(T): diagonals of T00 (lanes 0..15) and T11 (lanes 16..31), rows a < N
      double dr = 0.0, di = 0.0;
      if (lane < N && lane < kRows) {
        const double* tile = T + (lane < 16 ? 0 : 2) * kTile;
        const int x = lane & 15;
        dr = tile[tpos(x, x)];
        di = tile[tpos(x, x) + 1];
      }
      dr = row_sum16(dr);
      di = row_sum16(di);
      const double kr = readlane_d(dr, 15) + readlane_d(dr, 31);
      const double ki = readlane_d(di, 15) + readlane_d(di, 31);
      if (lane == 0 && active) {
        Kl[2 * ((long)w * K + n)] = kr;
        Kl[2 * ((long)w * K + n) + 1] = ki;
      }
    }
  };
  for (int n = nbeg; n < nend; n += 2) {
    vector(n, sB0, sU0, sB1);
    if (n + 1 < nend) vector(n + 1, sB1, sU1, sB0);
  }

  for (int off = 32; off > 0; off >>= 1) {
    exx_re += __shfl_xor(exx_re, off, 64);
    exx_im += __shfl_xor(exx_im, off, 64);
  }
  if (lane == 0 && active) {
    exx_part[2 * ((long)w * nslots + 2 * c + ta)] = exx_re;
    exx_part[2 * ((long)w * nslots + 2 * c + ta) + 1] = exx_im;
  }
}

// per Cholesky vector n: vmax[n] = max |component| of Lnak_n (one workgroup per vector)
__global__ void k_vector_max(long per, const double* __restrict__ L, double* __restrict__ vmax) {
  const long n = blockIdx.x;
  double m = 0.0;
  for (long e = threadIdx.x; e < 2 * per; e += blockDim.x) m = fmax(m, fabs(L[2 * n * per + e]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) vmax[n] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Lq[t][p][n][b][0..Mp) = slice t < S of the real (p = 0) or imaginary (p = 1) part of
// Lnak[n][b][:] / Lu[n][b] (zero for k >= M); one thread per row (n, b)
template <int S>
__global__ void k_split_lnak_i8(int rows, int M, int Mp, long KNMp, const double* __restrict__ L,
                                int8_t* __restrict__ Lq, double* __restrict__ Lu) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const double* x = L + 2 * (long)row * M;
  double mx = 0.0;
  bool bad = false;
  for (int k = 0; k < 2 * M; ++k) {
    if (!(fabs(x[k]) <= 1.0e300)) bad = true;
    mx = fmax(mx, fabs(x[k]));
  }
  bool zero;
  const double u = row_unit(bad ? INFINITY : mx, zero);
  const double inv = (u == u && !zero) ? 1.0 / u : 0.0;
  Lu[row] = u;
  for (int k = 0; k < Mp; ++k) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      int q[S];
      slice<S>(k < M ? x[2 * k + p] * inv : 0.0, q);
#pragma unroll
      for (int t = 0; t < S; ++t) Lq[(t * 2 + p) * KNMp + (long)row * Mp + k] = (int8_t)q[t];
    }
  }
}

}  // namespace

bool exchange_i8_supported(int M, int N) { return N <= 32 && M <= 64 * kMaxKC; }

int exchange_i8_row(int M) { return (M + 15) / 16 * 16; }

// slices of each Lnak row: 5 for up to 16 electrons, 6 for 17..32 (per-vector orders)
int exchange_i8_slices(int N) { return N > 16 ? kS6 : kSL; }

size_t exchange_i8_bytes(int M, int N, int K) {
  return (size_t)exchange_i8_slices(N) * 2 * K * N * exchange_i8_row(M);
}

void exchange_i8_orders(hipStream_t s, int M, int N, int K, const double* Lnak, int* ord) {
  double* vmax = nullptr;
  AFQ_HIP(hipMalloc(&vmax, sizeof(double) * std::max(K, 1)));
  std::vector<double> h(K);
  try {
    hipLaunchKernelGGL(k_vector_max, dim3(K), dim3(256), 0, s, (long)N * M, Lnak, vmax);
    AFQ_HIP(hipGetLastError());
    AFQ_HIP(hipMemcpyAsync(h.data(), vmax, sizeof(double) * K, hipMemcpyDeviceToHost, s));
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(vmax);
    throw;
  }
  AFQ_HIP(hipFree(vmax));
  double top = 0.0;
  for (double v : h) top = std::max(top, v);
  std::vector<int> o(K, kO6 - 1);
  for (int n = 0; n < K; ++n) {
    if (!(top > 0.0) || !(h[n] <= top)) continue;   // zero or non-finite: keep every product
    const double r = h[n] / top;
    int on = r > 0.0 ? (int)std::ceil(5.0 + 2.0 * std::log(r) / std::log(kBase)) : 2;
    o[n] = std::min(kO6 - 1, std::max(2, on));
  }
  AFQ_HIP(hipMemcpyAsync(ord, o.data(), sizeof(int) * K, hipMemcpyHostToDevice, s));
  AFQ_HIP(hipStreamSynchronize(s));
}

// whole rounds of the resident workgroups (two per CU), in multiples of the staging group; for
// N > 16 (two walkers per workgroup, one vector staged at a time) chunks of at least 2 vectors,
// so that the 2 tr(T^2) partials per chunk fit the K partials per walker of the workspace
int exchange_i8_chunk(int W, int N, int K) {
  const int step = N > 16 ? 2 : stage_of(1);
  const int wpb = N > 16 ? kT2Walkers : kWPB;
  const int cus = device_cus() * (N > 16 ? 1 : kWgPerCu);
  const int wgroups = (W + wpb - 1) / wpb;
  // N > 16: every chunk re-reads and re-slices G of its walkers, so the chunks may span all
  // vectors (one workgroup round when the walker groups fill the GPU)
  const int jmax = N > 16 ? (K + step - 1) / step : 16;
  int best = step;
  double best_eff = -1.0;
  for (int j = 1; j <= jmax; ++j) {
    const int chunk = step * j;
    const long nwg = (long)wgroups * ((K + chunk - 1) / chunk);
    const long rounds = (nwg + cus - 1) / cus;
    const double eff = (double)wgroups * K / ((double)rounds * cus * chunk);
    if (eff > best_eff + 1e-9 || (eff > best_eff - 1e-9 && chunk > best)) {
      best_eff = eff;
      best = chunk;
    }
  }
  return best;
}

// tr(T^2) partials per walker: one per chunk, two (one per row tile) for N > 16
int exchange_i8_chunks(int W, int N, int K) {
  const int chunk = exchange_i8_chunk(W, N, K);
  return ((K + chunk - 1) / chunk) * (N > 16 ? 2 : 1);
}

void split_lnak_i8(hipStream_t s, int M, int N, int K, const double* Lnak, int8_t* Lq,
                   double* Lu) {
  const int rows = K * N;
  const int Mp = exchange_i8_row(M);
  if (N > 16)
    hipLaunchKernelGGL(k_split_lnak_i8<kS6>, dim3((unsigned)((rows + 127) / 128)), dim3(128), 0, s,
                       rows, M, Mp, (long)K * N * Mp, Lnak, Lq, Lu);
  else
    hipLaunchKernelGGL(k_split_lnak_i8<kSL>, dim3((unsigned)((rows + 127) / 128)), dim3(128), 0, s,
                       rows, M, Mp, (long)K * N * Mp, Lnak, Lq, Lu);
  AFQ_HIP(hipGetLastError());
}

void launch_exchange_i8(hipStream_t s, int W, int M, int N, int K, const double* G,
                        const int8_t* Lq, const double* Lu, const int* ord, double* Kl,
                        double* exx_part) {
  if (W <= 0 || K <= 0) return;
  AFQ_CHECK(exchange_i8_supported(M, N), AFQ_EUNSUPPORTED,
            "int8 exchange: needs N <= 32 and M <= 128");
  AFQ_CHECK(exchange_i8_bytes(M, N, K) < (1UL << 31), AFQ_EUNSUPPORTED,
            "int8 exchange: the Cholesky slices exceed 2 GiB");
  AFQ_CHECK(N <= 16 || ord != nullptr, AFQ_EINVAL, "int8 exchange: no slice orders");
  const int chunk = exchange_i8_chunk(W, N, K);
  const int nchunk = (K + chunk - 1) / chunk;
  const int Mp = exchange_i8_row(M);
  if (N > 16) {
    dim3 grid((W + kT2Walkers - 1) / kT2Walkers, nchunk);
    const int nslots = 2 * nchunk;
    if (M <= 64)
      hipLaunchKernelGGL((k_exchange_i8_t2<1>), grid, dim3(128 * kT2Walkers), 0, s, W, M, N, K,
                         Mp, chunk, G, Lq, Lu, ord, Kl, exx_part, nslots);
    else
      hipLaunchKernelGGL((k_exchange_i8_t2<2>), grid, dim3(128 * kT2Walkers), 0, s, W, M, N, K,
                         Mp, chunk, G, Lq, Lu, ord, Kl, exx_part, nslots);
  } else {
    dim3 grid((W + kWPB - 1) / kWPB, nchunk);
    if (M <= 64)
      hipLaunchKernelGGL((k_exchange_i8<1, 2>), grid, dim3(64 * kWPB), 0, s, W, M, N, K, Mp,
                         chunk, G, Lq, Lu, Kl, exx_part, nchunk);
    else
      hipLaunchKernelGGL((k_exchange_i8<2, 1>), grid, dim3(64 * kWPB), 0, s, W, M, N, K, Mp,
                         chunk, G, Lq, Lu, Kl, exx_part, nchunk);
  }
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
