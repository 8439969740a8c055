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

// orders LDS traffic of one wavefront: its own writes are visible to its own later reads
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
  __shared__ double sT[kWPB][16 * 16 * 2];          // per-wavefront T block [a][b] (re,im)
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

  // group loads: slice rows of vectors nv0 .. nv0 + kStage - 1 into registers (zero outside)
  uint4 pre[kPre];
  double preU = 0.0;
  auto load_group = [&](int nv0) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int e = threadIdx.x + i * 64 * kWPB;
      const int kv = e % nvec;
      const int b = (e / nvec) % 16;
      const int tp = (e / (nvec * 16)) % (kSL * 2);
      const int v = e / (nvec * 16 * kSL * 2);
      const int n = nv0 + v;
      pre[i] = uint4{0u, 0u, 0u, 0u};
      if (e < kGroupVec && n < nend && b < N && 16 * kv < Mp)
        pre[i] = *reinterpret_cast<const uint4*>(Lq + tp * KNMp + ((long)n * N + b) * Mp + 16 * kv);
    }
    if (threadIdx.x < kStage * 16) {
      const int v = threadIdx.x / 16, b = threadIdx.x % 16, n = nv0 + v;
      preU = (n < nend && b < N) ? Lu[(long)n * N + b] : 0.0;
    }
  };
  auto store_group = [&](int buf) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int e = threadIdx.x + i * 64 * kWPB;
      const int kv = e % nvec;
      const int b = (e / nvec) % 16;
      const int tp = (e / (nvec * 16)) % (kSL * 2);
      const int v = e / (nvec * 16 * kSL * 2);
      if (e < kGroupVec)
        *reinterpret_cast<uint4*>(sBuf[buf] + ((v * kSL * 2 + tp) * 16 + b) * ROWB + 16 * kv) =
            pre[i];
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
            const int pr01 = 254 * accR[v][0][r] + accR[v][1][r];
            const int pi01 = 254 * accI[v][0][r] + accI[v][1][r];
            const int pr23 = 254 * accR[v][2][r] + accR[v][3][r];
            const int pi23 = 254 * accI[v][2][r] + accI[v][3][r];
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
          T[2 * (rw * 16 + r16)] = tr[r];
          T[2 * (rw * 16 + r16) + 1] = ti[r];
        }
        wave_lds_sync();
        if (n < nend) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rw = acc_row<float>(g, r);
            const double yr = T[2 * (r16 * 16 + rw)], yi = T[2 * (r16 * 16 + rw) + 1];
            exx_re += tr[r] * yr - ti[r] * yi;
            exx_im += tr[r] * yi + ti[r] * yr;
          }
          // tr(T): the N diagonal entries, summed over lanes 0..15 in four rounds
          double dr = 0.0, di = 0.0;
          if (lane < N) {
            dr = T[2 * (lane * 16 + lane)];
            di = T[2 * (lane * 16 + lane) + 1];
          }
          for (int off = 8; off > 0; off >>= 1) {
            dr += __shfl_xor(dr, off, 16);
            di += __shfl_xor(di, off, 16);
          }
          if (lane == 0 && active) {
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

// Lq[t][p][n][b][0..Mp) = slice t of the real (p = 0) or imaginary (p = 1) part of
// Lnak[n][b][:] / Lu[n][b] (zero for k >= M); one thread per row (n, b)
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
      int q[kSL];
      slice<kSL>(k < M ? x[2 * k + p] * inv : 0.0, q);
#pragma unroll
      for (int t = 0; t < kSL; ++t) Lq[(t * 2 + p) * KNMp + (long)row * Mp + k] = (int8_t)q[t];
    }
  }
}

}  // namespace

bool exchange_i8_supported(int M, int N) { return N <= 16 && M <= 64 * kMaxKC; }

int exchange_i8_row(int M) { return (M + 15) / 16 * 16; }

size_t exchange_i8_bytes(int M, int N, int K) {
  return (size_t)kSL * 2 * K * N * exchange_i8_row(M);
}

// whole rounds of one workgroup per CU, in multiples of the largest staging group
int exchange_i8_chunk(int W, int K) {
  constexpr int kStage = stage_of(1);
  const int cus = device_cus() * kWgPerCu;
  const int wgroups = (W + kWPB - 1) / kWPB;
  int best = kStage;
  double best_eff = -1.0;
  for (int j = 1; j <= 8; ++j) {
    const int chunk = kStage * j;
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

int exchange_i8_chunks(int W, int K) {
  const int chunk = exchange_i8_chunk(W, K);
  return (K + chunk - 1) / chunk;
}

void split_lnak_i8(hipStream_t s, int M, int N, int K, const double* Lnak, int8_t* Lq,
                   double* Lu) {
  const int rows = K * N;
  const int Mp = exchange_i8_row(M);
  hipLaunchKernelGGL(k_split_lnak_i8, dim3((unsigned)((rows + 127) / 128)), dim3(128), 0, s, rows,
                     M, Mp, (long)K * N * Mp, Lnak, Lq, Lu);
  AFQ_HIP(hipGetLastError());
}

void launch_exchange_i8(hipStream_t s, int W, int M, int N, int K, const double* G,
                        const int8_t* Lq, const double* Lu, double* Kl, double* exx_part) {
  if (W <= 0 || K <= 0) return;
  AFQ_CHECK(exchange_i8_supported(M, N), AFQ_EUNSUPPORTED,
            "int8 exchange: needs N <= 16 and M <= 128");
  const int chunk = exchange_i8_chunk(W, K);
  const int nchunk = (K + chunk - 1) / chunk;
  const int Mp = exchange_i8_row(M);
  dim3 grid((W + kWPB - 1) / kWPB, nchunk);
  if (M <= 64)
    hipLaunchKernelGGL((k_exchange_i8<1, 2>), grid, dim3(64 * kWPB), 0, s, W, M, N, K, Mp, chunk,
                       G, Lq, Lu, Kl, exx_part, nchunk);
  else
    hipLaunchKernelGGL((k_exchange_i8<2, 1>), grid, dim3(64 * kWPB), 0, s, W, M, N, K, Mp, chunk,
                       G, Lq, Lu, Kl, exx_part, nchunk);
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
