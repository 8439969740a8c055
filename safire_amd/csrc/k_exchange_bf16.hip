// Exchange and Coulomb energy (Real3IndexFactorization::energy_impl, Real3IndexFactorization.hpp:
// 476-559) with the contraction T_wn = G_w Lnak_n^T on BF16 MFMA (v_mfma_f32_16x16x32_bf16),
// the BF16X3 form: three real products per complex product and hi/lo BF16 splits.
//
// Complex product (Gauss): with P1 = Re(a) Re(b), P2 = Im(a) Im(b), P3 = (Re a + Im a)(Re b + Im b),
// Re(ab) = P1 - P2 and Im(ab) = P3 - P1 - P2: three real products instead of four. Every real
// operand x is split into x = h + l, h = bf16(x), l = bf16(x - h) (~16 significant bits), and a
// real product is h h' + h l' + l h' (the l l' term, ~2^-16 of the product, is dropped): nine
// BF16 MFMAs per 16x16x32 complex tile, FP32 accumulation. The sums Re + Im are split on their
// own (not summed from the parts), so every operand of an MFMA is a plain hi/lo pair.
// Accuracy: each operand to ~2^-16 relative, each T entry to ~2^-15 of its terms, so this form
// is ~100x less accurate than FP32 and far from a 1e-5 Ha energy gate on these Hamiltonians; it
// is the fastest form and reports its error next to the others.
//
// Structure (as the int8 exchange for 17..32 electrons): a workgroup of 8 wavefronts holds 8/NT
// walkers, NT = 1 or 2 wavefronts per walker (one per 16-row tile of T); each wavefront splits its
// rows of G once into MFMA A-fragments kept in registers; the six BF16 planes of one Cholesky
// vector (hi and lo of Re, Im and Re + Im, precomputed at set-up, rows padded to 16 NT and
// columns to 32 KC with zeros) are copied from global memory straight into LDS
// (global_load_lds_dwordx4), double-buffered in two LDS objects with the loop unrolled by two, so
// the copy of the next vector runs under this vector's MFMAs; each wavefront writes its rows of
// T (FP32) to the walker's LDS block, and tr(T^2) and tr(T) are summed in FP64 after one barrier.
// N <= 32, M <= 128.
#include "afq_internal.hpp"

namespace afq {

namespace {

constexpr int kPlanes = 6;       // hi, lo of Re, Im, Re + Im
constexpr int kWaves = 8;        // wavefronts per workgroup

// 16-byte chunk position of logical chunk c in row b of a staged plane: spreads the 16 rows an
// MFMA fragment reads over distinct bank groups
template <int KC>
__device__ inline int bswz(int b) {
  return KC == 4 ? (b & 15) : KC == 2 ? ((b >> 1) & 7) : ((b >> 2) & 3);
}

__device__ inline void split2(double x, __bf16& h, __bf16& l) {
  h = (__bf16)(float)x;
  l = (__bf16)(float)(x - (double)(float)h);
}

// workgroup barrier after this wavefront's LDS accesses completed; copies into LDS in flight are
// not waited for (they are waited for explicitly before the buffer they fill is read)
__device__ inline void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ inline f4 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f4 c) {
  c = mfma_bf16(ah, bh, c);
  c = mfma_bf16(ah, bl, c);
  return mfma_bf16(al, bh, c);
}

template <int KC, int NT>
__global__ __launch_bounds__(64 * kWaves, 1) void k_exchange_bf16(
    int W, int M, int N, int K, int kChunk, const double* __restrict__ G,
    const __bf16* __restrict__ Lb, double* __restrict__ Kl, double* __restrict__ exx_part,
    int nslots) {
  constexpr int kRows = 16 * NT;                  // rows (and columns) of a T block
  constexpr int kWalkers = kWaves / NT;
  constexpr int Mp = 32 * KC;
  constexpr int ROWB = 2 * Mp;                    // bytes of a staged row
  constexpr int kPlane = kRows * ROWB;
  constexpr int kVecBytes = kPlanes * kPlane;
  constexpr int kWindows = kVecBytes / 1024;      // 1 KB copied per wavefront instruction
  static_assert(kVecBytes % 1024 == 0, "a vector's planes are whole 1 KB windows");
  __shared__ __attribute__((aligned(16))) int8_t sB0[kVecBytes];
  __shared__ __attribute__((aligned(16))) int8_t sB1[kVecBytes];
  // per walker T [a][b] (re, im) in FP32; entry (x, y) sits in column y ^ (x % 16) of row x, so
  // that 16 lanes reading down a column hit distinct banks
  __shared__ float sT[kWalkers][kRows * kRows * 2];
  auto tcol = [](int x, int y) { return y ^ (x & 15); };
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int wl = wave / NT, ta = wave % NT;
  const int w_raw = blockIdx.x * kWalkers + wl;
  const bool active = w_raw < W;
  const int w = active ? w_raw : W - 1;
  const int c = blockIdx.y;
  const int nbeg = c * kChunk;
  const int nend = min(K, nbeg + kChunk);
  const long vecElems = (long)kRows * Mp;         // bf16 of one plane of one vector
  const long planeElems = (long)K * vecElems;
  const int arow = ta * 16 + r16;                   // this lane's row of G

  // A fragments of G_w rows 16 ta .. 16 ta + 15: hi/lo of Re, Im, Re + Im; k = 32 kc + 8 g + q
  bf16x8 frag[KC][kPlanes];
  {
    const double* Gw = G + 2 * ((long)w * N + arow) * M;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = 32 * kc + 8 * g + q;
        double xr = 0.0, xi = 0.0;
        if (arow < N && k < M) {
          xr = Gw[2 * k];
          xi = Gw[2 * k + 1];
        }
        __bf16 h, l;
        split2(xr, h, l);
        frag[kc][0][q] = h;
        frag[kc][1][q] = l;
        split2(xi, h, l);
        frag[kc][2][q] = h;
        frag[kc][3][q] = l;
        split2(xr + xi, h, l);
        frag[kc][4][q] = h;
        frag[kc][5][q] = l;
      }
  }

  double exx_re = 0.0, exx_im = 0.0;
  float* T = sT[wl];

  // copy vector n's planes into dst: window win of the vector is LDS bytes
  // [1024 win, 1024 win + 1024), lane l writes bytes 16 l .. 16 l + 15 of it
  auto stage = [&](int n, int8_t* dst) {
    for (int win = wave; win < kWindows; win += kWaves) {
      const int o = win * 1024 + 16 * lane;
      const int p = o / kPlane;
      const int b = (o % kPlane) / ROWB;
      const int kv = ((o % ROWB) >> 4) ^ bswz<KC>(b);
      const __bf16* src = Lb + p * planeElems + (long)n * vecElems + (long)b * Mp + 8 * kv;
      __builtin_amdgcn_global_load_lds(
          (const void __attribute__((address_space(1)))*)src,
          (void __attribute__((address_space(3)))*)(dst + win * 1024), 16, 0, 0);
    }
  };

  if (nbeg < nend) stage(nbeg, sB0);
  auto vector = [&](int n, const int8_t* sBb, int8_t* sBn) {
    // vector n is in sBb (this thread's copies done, then everyone's); every wavefront is past
    // vector n - 1, so sBn and the T blocks are free again
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    __syncthreads();
    if (n + 1 < nend) stage(n + 1, sBn);
    // both column tiles together: P1, P2, P3 of tile tb
    f4 acc[NT][3];
#pragma unroll
    for (int tb = 0; tb < NT; ++tb)
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[tb][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      bf16x8 bf[NT][kPlanes];
#pragma unroll
      for (int tb = 0; tb < NT; ++tb) {
        const int row = tb * 16 + r16;
        const int chunk = ((4 * kc + g) ^ bswz<KC>(row)) << 4;
#pragma unroll
        for (int p = 0; p < kPlanes; ++p)
          bf[tb][p] = *reinterpret_cast<const bf16x8*>(sBb + p * kPlane + row * ROWB + chunk);
      }
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int tb = 0; tb < NT; ++tb)
          acc[tb][j] = mfma3(frag[kc][2 * j], frag[kc][2 * j + 1], bf[tb][2 * j],
                             bf[tb][2 * j + 1], acc[tb][j]);
    }
    // T entries of this wavefront's rows: Re = P1 - P2, Im = P3 - P1 - P2 (rows 16 ta + 4 g + r,
    // columns 16 tb + r16)
    float tre[NT][4], tim[NT][4];
#pragma unroll
    for (int tb = 0; tb < NT; ++tb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float p1 = acc[tb][0][r], p2 = acc[tb][1][r], p3 = acc[tb][2][r];
        tre[tb][r] = p1 - p2;
        tim[tb][r] = p3 - p1 - p2;
        const int a = ta * 16 + 4 * g + r, b = tb * 16 + r16;
        T[2 * (a * kRows + tcol(a, b))] = tre[tb][r];
        T[2 * (a * kRows + tcol(a, b)) + 1] = tim[tb][r];
      }
    lds_barrier();   // every row tile of the walker's T is in LDS
#pragma unroll
    for (int tb = 0; tb < NT; ++tb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = ta * 16 + 4 * g + r, b = tb * 16 + r16;
        const double xr = tre[tb][r], xi = tim[tb][r];
        const double yr = T[2 * (b * kRows + tcol(b, a))];
        const double yi = T[2 * (b * kRows + tcol(b, a)) + 1];
        exx_re += xr * yr - xi * yi;
        exx_im += xr * yi + xi * yr;
      }
    if (ta == 0) {
      // tr(T): diagonal entries a < N in lanes 0 .. kRows - 1
      double dr = 0.0, di = 0.0;
      if (lane < N && lane < kRows) {
        dr = T[2 * (lane * kRows + tcol(lane, lane))];
        di = T[2 * (lane * kRows + tcol(lane, lane)) + 1];
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
  };
  for (int n = nbeg; n < nend; n += 2) {
    vector(n, sB0, sB1);
    if (n + 1 < nend) vector(n + 1, sB1, sB0);
  }

  for (int off = 32; off > 0; off >>= 1) {
    exx_re += __shfl_xor(exx_re, off, 64);
    exx_im += __shfl_xor(exx_im, off, 64);
  }
  if (lane == 0 && active) {
    exx_part[2 * ((long)w * nslots + NT * c + ta)] = exx_re;
    exx_part[2 * ((long)w * nslots + NT * c + ta) + 1] = exx_im;
  }
}

// Lb[p][n][b][k], b < rows, k < Mp: the six BF16 planes of Lnak[n][b][k] (zero for b >= N or
// k >= M): hi, lo of Re, Im and Re + Im. One thread per (n, b, k).
__global__ void k_split_lnak_bf16(int K, int N, int M, int rows, int Mp,
                                  const double* __restrict__ L, __bf16* __restrict__ Lb) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)rows * Mp;
  if (e >= (long)K * per) return;
  const long n = e / per;
  const int b = (int)((e % per) / Mp), k = (int)(e % Mp);
  double xr = 0.0, xi = 0.0;
  if (b < N && k < M) {
    xr = L[2 * ((n * N + b) * M + k)];
    xi = L[2 * ((n * N + b) * M + k) + 1];
  }
  const long plane = (long)K * per;
  __bf16 h, l;
  split2(xr, h, l);
  Lb[0 * plane + e] = h;
  Lb[1 * plane + e] = l;
  split2(xi, h, l);
  Lb[2 * plane + e] = h;
  Lb[3 * plane + e] = l;
  split2(xr + xi, h, l);
  Lb[4 * plane + e] = h;
  Lb[5 * plane + e] = l;
}

int kc_of(int M) { return M <= 32 ? 1 : M <= 64 ? 2 : 4; }
int nt_of(int N) { return N <= 16 ? 1 : 2; }

}  // namespace

bool exchange_bf16_supported(int M, int N) { return N <= 32 && M <= 128; }

size_t exchange_bf16_elems(int M, int N, int K) {
  return (size_t)kPlanes * K * 16 * nt_of(N) * 32 * kc_of(M);
}

// whole rounds of one workgroup per CU, chunks of an even number of vectors (the loop is
// unrolled by two) with at most K + 1 tr(T^2) partials per walker
int exchange_bf16_chunk(int W, int N, int K) {
  const int cus = device_cus();
  const int wpb = kWaves / nt_of(N);
  const int wgroups = (W + wpb - 1) / wpb;
  int best = 2;
  double best_eff = -1.0;
  for (int j = 1; j <= 16; ++j) {
    const int chunk = 2 * j;
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

int exchange_bf16_chunks(int W, int N, int K) {
  const int chunk = exchange_bf16_chunk(W, N, K);
  return ((K + chunk - 1) / chunk) * nt_of(N);
}

void split_lnak_bf16(hipStream_t s, int M, int N, int K, const double* Lnak, __bf16* Lb) {
  const int rows = 16 * nt_of(N), Mp = 32 * kc_of(M);
  const long tot = (long)K * rows * Mp;
  hipLaunchKernelGGL(k_split_lnak_bf16, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, K,
                     N, M, rows, Mp, Lnak, Lb);
  AFQ_HIP(hipGetLastError());
}

void launch_exchange_bf16(hipStream_t s, int W, int M, int N, int K, const double* G,
                          const __bf16* Lb, double* Kl, double* exx_part) {
  if (W <= 0 || K <= 0) return;
  AFQ_CHECK(exchange_bf16_supported(M, N), AFQ_EUNSUPPORTED,
            "BF16 exchange: needs N <= 32 and M <= 128");
  const int chunk = exchange_bf16_chunk(W, N, K);
  const int nchunk = (K + chunk - 1) / chunk;
  const int nt = nt_of(N), kc = kc_of(M);
  const int nslots = nchunk * nt;
  dim3 grid((W + kWaves / nt - 1) / (kWaves / nt), nchunk);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(64 * kWaves), 0, s, W, M, N, K, chunk, G, Lb, Kl,
                       exx_part, nslots);
  };
  if (nt == 1) {
    if (kc == 1) go(k_exchange_bf16<1, 1>);
    else if (kc == 2) go(k_exchange_bf16<2, 1>);
    else go(k_exchange_bf16<4, 1>);
  } else {
    if (kc == 1) go(k_exchange_bf16<1, 2>);
    else if (kc == 2) go(k_exchange_bf16<2, 2>);
    else go(k_exchange_bf16<4, 2>);
  }
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
