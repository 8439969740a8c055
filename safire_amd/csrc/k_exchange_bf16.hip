// Exchange and Coulomb energy (Real3IndexFactorization::energy_impl, Real3IndexFactorization.hpp:
// 476-559) with the contraction T_wn = G_w Lnak_n^T on BF16 MFMA (v_mfma_f32_16x16x32_bf16).
//
// Both operands are split into three BF16 slices, x = x0 + x1 + x2 (hi, mid, lo: ~24 bits, the
// remainder below 2^-24 |x|), and the six slice products with i + j <= 2 are accumulated in
// FP32. Complex products take 4 real products (the real part uses -Im(G), negated exactly in
// BF16). Accuracy: the slices represent each operand to ~2^-24, like FP32 rounding, and the FP32
// accumulator rounds each T entry to ~2^-24 again; the rounding of G is shared by all Cholesky
// vectors and adds up coherently in EXX. On walkers whose |EXX| reaches 1e5 Ha the energy error
// is ~1e-4 Ha, within 3x of the FP32 form: this form does not meet a 1e-5 Ha gate there (the
// int8 Ozaki split of k_exchange_i8.hip does).
//
// Structure follows the FP64 kernel: a workgroup holds WPB walkers (one per wavefront) and a
// chunk of Cholesky vectors; each wavefront splits its walker's G once into MFMA A-fragments kept
// in registers; the NV vectors of a group are staged in LDS as 3 slices x (re, im) of Lnak; every
// N x N block of T is transposed through per-wavefront LDS for tr(T^2) and tr(T) in FP64.
// N <= 16 (one 16x16 tile per block), M <= 96.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kNV = 4;     // vectors staged per group
constexpr int kWPB = 8;    // walkers (wavefronts) per workgroup
constexpr int kMaxKC = 3;  // 32-wide k chunks: M <= 96

__device__ inline void split3(double x, __bf16& a, __bf16& b, __bf16& c) {
  a = (__bf16)(float)x;
  const double r1 = x - (double)(float)a;
  b = (__bf16)(float)r1;
  const double r2 = r1 - (double)(float)b;
  c = (__bf16)(float)r2;
}

__device__ inline bf16x8 bneg(bf16x8 v) {
  bf16x8 o;
#pragma unroll
  for (int q = 0; q < 8; ++q) o[q] = -v[q];
  return o;
}

// LDS row stride (bf16) of a staged Lnak row: M padded to a multiple of 32 plus 8, so that the
// 16 rows read by a quarter-wavefront fall into different banks
__host__ __device__ inline int lds_row(int M) { return ((M + 31) / 32) * 32 + 8; }

template <int KC>
__global__ __launch_bounds__(64 * kWPB) void k_exchange_bf16(
    int W, int M, int N, int K, int kChunk, const double* __restrict__ G,
    const __bf16* __restrict__ Lb, double* __restrict__ Kl, double* __restrict__ exx_part,
    int nchunk) {
  extern __shared__ __bf16 sB[];                    // [kNV][3][2][16][row]
  // per-wavefront T block [b][a] (re, im), rows padded to 17 entries (distinct banks for the
  // transposed read)
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
  const int row = lds_row(M);
  const long KNM = (long)K * N * M;

  // A fragments of G_w: [kc][slice] for Re and Im, row a = r16, k = 32 kc + 8 g + q
  bf16x8 ar[KC][3], ai[KC][3];
  {
    const double* Gw = G + 2 * (long)w * N * M;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = 32 * kc + 8 * g + q;
        double xr = 0.0, xi = 0.0;
        if (r16 < N && k < M) {
          xr = Gw[2 * ((long)r16 * M + k)];
          xi = Gw[2 * ((long)r16 * M + k) + 1];
        }
        __bf16 h, m, l;
        split3(xr, h, m, l);
        ar[kc][0][q] = h;
        ar[kc][1][q] = m;
        ar[kc][2][q] = l;
        split3(xi, h, m, l);
        ai[kc][0][q] = h;
        ai[kc][1][q] = m;
        ai[kc][2][q] = l;
      }
    }
  }

  double exx_re = 0.0, exx_im = 0.0;
  double* T = sT[wave];

  for (int nv0 = nbeg; nv0 < nend; nv0 += kNV) {
    __syncthreads();
    // stage slices s, parts p of Lnak[nv0 + v][a][0..M) for v < kNV, a < 16
    const int kvec = M / 8;   // 16-byte vectors per row (M % 8 == 0 checked by the launcher)
    for (int e = threadIdx.x; e < kNV * 6 * 16 * kvec; e += 64 * kWPB) {
      const int kv = e % kvec;
      const int a = (e / kvec) % 16;
      const int sp = (e / (kvec * 16)) % 6;
      const int v = e / (kvec * 16 * 6);
      const int n = nv0 + v;
      uint4 val = {0, 0, 0, 0};
      if (n < nend && a < N)
        val = *reinterpret_cast<const uint4*>(Lb + sp * KNM + ((long)n * N + a) * M + 8 * kv);
      *reinterpret_cast<uint4*>(sB + (((v * 6 + sp) * 16 + a) * row) + 8 * kv) = val;
    }
    // zero the padding k >= M of each row up to the 32-wide chunk
    const int kpad = KC * 32;
    if (kpad > M) {
      for (int e = threadIdx.x; e < kNV * 6 * 16 * (kpad - M); e += 64 * kWPB) {
        const int k = M + e % (kpad - M);
        const int rr = e / (kpad - M);
        sB[rr * row + k] = (__bf16)0.0f;
      }
    }
    __syncthreads();

    f4 tr[kNV], ti[kNV];
#pragma unroll
    for (int v = 0; v < kNV; ++v) {
      tr[v] = f4{0, 0, 0, 0};
      ti[v] = f4{0, 0, 0, 0};
    }
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
      for (int v = 0; v < kNV; ++v) {
        bf16x8 br[3], bi[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          br[s] = *reinterpret_cast<const bf16x8*>(sB + ((v * 6 + 2 * s) * 16 + r16) * row +
                                                   32 * kc + 8 * g);
          bi[s] = *reinterpret_cast<const bf16x8*>(sB + ((v * 6 + 2 * s + 1) * 16 + r16) * row +
                                                   32 * kc + 8 * g);
        }
        // slice products (i, j) with i + j <= 2
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const bf16x8 nai = bneg(ai[kc][i]);
#pragma unroll
          for (int j = 0; j + i <= 2; ++j) {
            tr[v] = mfma_bf16(ar[kc][i], br[j], tr[v]);
            tr[v] = mfma_bf16(nai, bi[j], tr[v]);
            ti[v] = mfma_bf16(ar[kc][i], bi[j], ti[v]);
            ti[v] = mfma_bf16(ai[kc][i], br[j], ti[v]);
          }
        }
      }
    }

    // epilogue in FP64: T block -> LDS, tr(T^2) lane partials, tr(T) -> Kl
#pragma unroll
    for (int v = 0; v < kNV; ++v) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rw = acc_row<float>(g, r), cl = r16;
        *reinterpret_cast<double2*>(T + 2 * (rw * 17 + cl)) =
            double2{(double)tr[v][r], (double)ti[v][r]};
      }
      __syncthreads();
      const int n = nv0 + v;
      if (n < nend) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rw = acc_row<float>(g, r), cl = r16;
          const double xr = (double)tr[v][r], xi = (double)ti[v][r];
          const double2 y = *reinterpret_cast<const double2*>(T + 2 * (cl * 17 + rw));
          const double yr = y.x, yi = y.y;
          exx_re += xr * yr - xi * yi;
          exx_im += xr * yi + xi * yr;
        }
        double dr = 0.0, di = 0.0;
        if (lane < N) {
          const double2 d = *reinterpret_cast<const double2*>(T + 2 * (lane * 17 + lane));
          dr = d.x;
          di = d.y;
        }
        for (int off = 32; off > 0; off >>= 1) {
          dr += __shfl_xor(dr, off, 64);
          di += __shfl_xor(di, off, 64);
        }
        if (lane == 0 && active) {
          Kl[2 * ((long)w * K + n)] = dr;
          Kl[2 * ((long)w * K + n) + 1] = di;
        }
      }
    }
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

// Lb[s][p][n][a][k] = slice s of the real (p = 0) or imaginary (p = 1) part of Lnak[n][a][k]
__global__ void k_split_lnak(long KNM, const double* __restrict__ L, __bf16* __restrict__ Lb) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= KNM) return;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    __bf16 h, m, l;
    split3(L[2 * e + p], h, m, l);
    Lb[(0 * 2 + p) * KNM + e] = h;
    Lb[(1 * 2 + p) * KNM + e] = m;
    Lb[(2 * 2 + p) * KNM + e] = l;
  }
}

}  // namespace

bool exchange_bf16_supported(int M, int N) { return N <= 16 && M <= 32 * kMaxKC && M % 8 == 0; }

// same balancing rule as the FP64 kernel: whole rounds of one workgroup per CU
int exchange_bf16_chunk(int W, int K) {
  const int cus = device_cus();
  const int wgroups = (W + kWPB - 1) / kWPB;
  int best = kNV;
  double best_eff = -1.0;
  for (int j = 1; j <= 8; ++j) {
    const int chunk = kNV * j;
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

int exchange_bf16_chunks(int W, int K) {
  const int chunk = exchange_bf16_chunk(W, K);
  return (K + chunk - 1) / chunk;
}

void split_lnak_bf16(hipStream_t s, int M, int N, int K, const double* Lnak, __bf16* Lb) {
  const long KNM = (long)K * N * M;
  hipLaunchKernelGGL(k_split_lnak, dim3((unsigned)((KNM + 255) / 256)), dim3(256), 0, s, KNM,
                     Lnak, Lb);
  AFQ_HIP(hipGetLastError());
}

void launch_exchange_bf16(hipStream_t s, int W, int M, int N, int K, const double* G,
                          const __bf16* Lb, double* Kl, double* exx_part) {
  if (W <= 0 || K <= 0) return;
  AFQ_CHECK(exchange_bf16_supported(M, N), AFQ_EUNSUPPORTED,
            "BF16 exchange: needs N <= 16 and M <= 96 with M % 8 == 0");
  const int chunk = exchange_bf16_chunk(W, K);
  const int nchunk = (K + chunk - 1) / chunk;
  const size_t lds = sizeof(__bf16) * (size_t)kNV * 6 * 16 * lds_row(M);
  dim3 grid((W + kWPB - 1) / kWPB, nchunk);
  auto go = [&](auto kern) {
    if (lds + sizeof(double) * kWPB * 512 > 64 * 1024)   // dynamic + the static T blocks
      ensure_dynamic_lds((const void*)kern, std::max(lds, (size_t)64 * 1024 + 1));
    hipLaunchKernelGGL(kern, grid, dim3(64 * kWPB), lds, s, W, M, N, K, chunk, G, Lb, Kl,
                       exx_part, nchunk);
  };
  const int kc = (M + 31) / 32;
  if (kc <= 1) go(k_exchange_bf16<1>);
  else if (kc == 2) go(k_exchange_bf16<2>);
  else go(k_exchange_bf16<3>);
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
