// Exchange and Coulomb energy of the Cholesky Hamiltonian (Real3IndexFactorization::energy_impl,
// Real3IndexFactorization.hpp:476-559), the dominant term of a local-energy step
// (8 N^2 M K flops per walker against ~50 N M K for everything else):
//
//   T_wn[b][a] = sum_k G_w[b][k] Lnak[n][a][k]        (N x N complex per walker w and vector n)
//   EXX_w      = -scl/2 sum_n sum_{a,b} T_wn[a][b] T_wn[b][a]   (non-conjugated)
//   Kl_wn      = sum_a T_wn[a][a],   EJ_w = (scl^2/2) sum_n Kl_wn^2     (scl = 2 closed shell)
//
// T is never written to memory. A workgroup holds WPB walkers (one per wavefront) and a chunk of
// Cholesky vectors; it walks the chunk NV vectors at a time. For each 16-wide slice of the
// orbital index k it stages the NV vectors' Lnak rows in LDS (shared by all wavefronts), each
// wavefront loads its walker's G slice straight from memory into MFMA A-fragments and
// accumulates NV x (NT x NT) complex 16x16 tiles (NT = ceil(N/16)). After the k loop every
// N x N block goes through a per-wavefront LDS scratch to read T^T, giving tr(T^2) and tr(T).
// The per-lane partial of tr(T^2) stays in registers across the whole chunk and is reduced once;
// Kl_wn is written per vector; a second kernel sums the chunk partials in a fixed order.
#include "afq_internal.hpp"

namespace afq {

namespace {

template <typename R, int NT, int NV, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_exchange(int W, int M, int N, int K, int kChunk,
                                                      const double* __restrict__ G,
                                                      const R* __restrict__ L,
                                                      double* __restrict__ Kl,
                                                      double* __restrict__ exx_part,
                                                      int nchunk) {
  typedef typename acc_of<R>::type acc_t;
  constexpr int NB = 16 * NT;
  constexpr int LDB = 17;                 // padded k stride of the staged Lnak slice
  __shared__ R sB[NV * NB * LDB * 2];     // [v][a][k] (re,im)
  // per-wavefront T block [b][a] (re, im), rows padded by one entry (distinct banks for the
  // transposed read)
  __shared__ __attribute__((aligned(16))) double sT[WPB][NB * (NB + 1) * 2];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, r16 = lane & 15;
  const int w_raw = blockIdx.x * WPB + wave;
  const bool active = w_raw < W;
  const int w = active ? w_raw : W - 1;
  const int c = blockIdx.y;
  const int nbeg = c * kChunk;
  const int nend = min(K, nbeg + kChunk);
  const double* Gw = G + 2 * (long)w * N * M;

  double exx_re = 0.0, exx_im = 0.0;

  for (int nv0 = nbeg; nv0 < nend; nv0 += NV) {
    acc_t tr_[NV][NT][NT], ti_[NV][NT][NT];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
          tr_[v][a][b] = acc_t{0, 0, 0, 0};
          ti_[v][a][b] = acc_t{0, 0, 0, 0};
        }

    for (int k0 = 0; k0 < M; k0 += 16) {
      __syncthreads();
      // stage Lnak[nv0+v][a][k0..k0+15] for v < NV, a < NB
      for (int e = threadIdx.x; e < NV * NB * 16; e += 64 * WPB) {
        const int kk = e & 15, va = e >> 4;
        const int a = va % NB, v = va / NB;
        const int n = nv0 + v, kx = k0 + kk;
        R re = 0, im = 0;
        if (n < nend && a < N && kx < M) {
          const R* p = L + 2 * (((long)n * N + a) * M + kx);
          re = p[0];
          im = p[1];
        }
        sB[2 * ((v * NB + a) * LDB + kk)] = re;
        sB[2 * ((v * NB + a) * LDB + kk) + 1] = im;
      }
      __syncthreads();
      // A fragments: G_w[t*16 + r16][k0 + 4g + s]   (16-byte loads when the slice is full)
      R ar[NT][4], ai[NT][4];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int row = t * 16 + r16;
        const int kb = k0 + 4 * g;
        if (row < N && kb + 4 <= M) {
          const double2* p = reinterpret_cast<const double2*>(Gw + 2 * ((long)row * M + kb));
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const double2 x = p[s];
            ar[t][s] = (R)x.x;
            ai[t][s] = (R)x.y;
          }
        } else {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const bool ok = row < N && kb + s < M;
            ar[t][s] = ok ? (R)Gw[2 * ((long)row * M + kb + s)] : (R)0;
            ai[t][s] = ok ? (R)Gw[2 * ((long)row * M + kb + s) + 1] : (R)0;
          }
        }
      }
      // k-step outer, vectors inner: consecutive MFMAs go to independent accumulators
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        R br[NV][NT], bi[NV][NT];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int u = 0; u < NT; ++u) {
            const int idx = 2 * ((v * NB + u * 16 + r16) * LDB + 4 * g + s);
            br[v][u] = sB[idx];
            bi[v][u] = sB[idx + 1];
          }
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b) {
              tr_[v][a][b] = mfma16x16x4(ar[a][s], br[v][b], tr_[v][a][b]);
              ti_[v][a][b] = mfma16x16x4(ar[a][s], bi[v][b], ti_[v][a][b]);
            }
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b) {
              tr_[v][a][b] = mfma16x16x4(-ai[a][s], bi[v][b], tr_[v][a][b]);
              ti_[v][a][b] = mfma16x16x4(ai[a][s], br[v][b], ti_[v][a][b]);
            }
      }
    }

    // epilogue: per vector, T block -> LDS, tr(T^2) lane partials, tr(T) -> Kl
    double* T = sT[wave];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      __syncthreads();
#pragma unroll
      for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = a * 16 + acc_row<R>(g, r), col = b * 16 + r16;
            *reinterpret_cast<double2*>(T + 2 * (row * (NB + 1) + col)) =
                double2{(double)tr_[v][a][b][r], (double)ti_[v][a][b][r]};
          }
      __syncthreads();
      const int n = nv0 + v;
      if (n < nend) {
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
          for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = a * 16 + acc_row<R>(g, r), col = b * 16 + r16;
              const double xr = (double)tr_[v][a][b][r], xi = (double)ti_[v][a][b][r];
              const double2 y =
                  *reinterpret_cast<const double2*>(T + 2 * (col * (NB + 1) + row));
              const double yr = y.x, yi = y.y;
              exx_re += xr * yr - xi * yi;
              exx_im += xr * yi + xi * yr;
            }
        double dr = 0.0, di = 0.0;
        if (lane < N) {
          const double2 d = *reinterpret_cast<const double2*>(T + 2 * (lane * (NB + 1) + lane));
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

// E[w][1] = -scl/2 sum_c exx_part[w][c];  E[w][2] = (scl^2/2) sum_n Kl[w][n]^2
__global__ void k_energy_finalize(int W, int K, int nchunk, double sx, double sj,
                                  const double* __restrict__ Kl,
                                  const double* __restrict__ exx_part, double* __restrict__ E) {
  const int w = blockIdx.x;
  const int lane = threadIdx.x;   // one wavefront
  double jr = 0, ji = 0;
  for (int n = lane; n < K; n += 64) {
    const double kr = Kl[2 * ((long)w * K + n)], ki = Kl[2 * ((long)w * K + n) + 1];
    jr += kr * kr - ki * ki;
    ji += 2 * kr * ki;
  }
  for (int off = 32; off > 0; off >>= 1) {
    jr += __shfl_xor(jr, off, 64);
    ji += __shfl_xor(ji, off, 64);
  }
  if (lane == 0) {
    double xr = 0, xi = 0;
    for (int c = 0; c < nchunk; ++c) {
      xr += exx_part[2 * ((long)w * nchunk + c)];
      xi += exx_part[2 * ((long)w * nchunk + c) + 1];
    }
    E[6 * (long)w + 2] = sx * xr;
    E[6 * (long)w + 3] = sx * xi;
    E[6 * (long)w + 4] = sj * jr;
    E[6 * (long)w + 5] = sj * ji;
  }
}

// E[w][0] = E0 + scl sum_{a,j} G[w][a][j] haj[a][j]  (Real3IndexFactorization.hpp:505-508)
__global__ __launch_bounds__(256) void k_e1(int NM, const double* __restrict__ G,
                                            const double* __restrict__ haj, double E0,
                                            double scl, double* __restrict__ E) {
  __shared__ double red[4][2];
  const int w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double sr = 0, si = 0;
  for (int e = tid; e < NM; e += 256) {
    const double gr = G[2 * ((long)w * NM + e)], gi = G[2 * ((long)w * NM + e) + 1];
    const double hr = haj[2 * e], hi = haj[2 * e + 1];
    sr += gr * hr - gi * hi;
    si += gr * hi + gi * hr;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  if (lane == 0) {
    red[wave][0] = sr;
    red[wave][1] = si;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0, b = 0;
    for (int q = 0; q < 4; ++q) {
      a += red[q][0];
      b += red[q][1];
    }
    E[6 * (long)w] = E0 + scl * a;
    E[6 * (long)w + 1] = scl * b;
  }
}

// Cholesky vectors per workgroup: a multiple of NV chosen so that the grid fills whole rounds
// of one workgroup per CU (a partial last round leaves CUs idle while it runs); among equally
// balanced choices the larger chunk (fewer workgroups, less per-workgroup setup) wins.
int choose_chunk(int W, int K, int NV, int WPB) {
  const int cus = device_cus();
  const int wgroups = (W + WPB - 1) / WPB;
  int best = NV;
  double best_eff = -1.0;
  for (int j = 1; j <= 8; ++j) {
    const int chunk = NV * j;
    const int nch = (K + chunk - 1) / chunk;
    const long nwg = (long)wgroups * nch;
    const long rounds = (nwg + cus - 1) / cus;
    // useful vector-work over the work of the rounds the grid occupies
    const double eff = (double)wgroups * K / ((double)rounds * cus * chunk);
    if (eff > best_eff + 1e-9 || (eff > best_eff - 1e-9 && chunk > best)) {
      best_eff = eff;
      best = chunk;
    }
  }
  return best;
}

constexpr int exx_nv(int N) { return N <= 16 ? 8 : 2; }
constexpr int exx_wpb(int N) { return N <= 16 ? 8 : 4; }

template <typename R, int NT, int NV, int WPB>
void launch_exchange_t(hipStream_t s, int W, int M, int N, int K, const double* G, const R* L,
                       double* Kl, double* exx_part) {
  const int chunk = choose_chunk(W, K, NV, WPB);
  const int nchunk = (K + chunk - 1) / chunk;
  dim3 grid((W + WPB - 1) / WPB, nchunk);
  hipLaunchKernelGGL((k_exchange<R, NT, NV, WPB>), grid, dim3(64 * WPB), 0, s, W, M, N, K, chunk,
                     G, L, Kl, exx_part, nchunk);
}

}  // namespace

int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  return cus;
}

int exx_nchunks(int W, int N, int K) {
  const int chunk = choose_chunk(W, K, exx_nv(N), exx_wpb(N));
  return (K + chunk - 1) / chunk;
}

void launch_exchange(hipStream_t s, int precision, int W, int M, int N, int K, const double* G,
                     const void* Lnak, double* Kl, double* exx_part) {
  if (W <= 0 || K <= 0) return;
  AFQ_CHECK(N <= 32, AFQ_EUNSUPPORTED, "exchange energy: more than 32 electrons per spin");
  if (precision == AFQ_FP64) {
    if (N <= 16)
      launch_exchange_t<double, 1, exx_nv(16), exx_wpb(16)>(s, W, M, N, K, G,
                                                            (const double*)Lnak, Kl, exx_part);
    else
      launch_exchange_t<double, 2, exx_nv(32), exx_wpb(32)>(s, W, M, N, K, G,
                                                            (const double*)Lnak, Kl, exx_part);
  } else {
    if (N <= 16)
      launch_exchange_t<float, 1, exx_nv(16), exx_wpb(16)>(s, W, M, N, K, G,
                                                           (const float*)Lnak, Kl, exx_part);
    else
      launch_exchange_t<float, 2, exx_nv(32), exx_wpb(32)>(s, W, M, N, K, G,
                                                           (const float*)Lnak, Kl, exx_part);
  }
  AFQ_HIP(hipGetLastError());
}

void launch_e1(hipStream_t s, int W, int M, int N, const double* G, const double* haj,
               double E0, double scl, double* E) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_e1, dim3(W), dim3(256), 0, s, N * M, G, haj, E0, scl, E);
  AFQ_HIP(hipGetLastError());
}

void launch_energy_finalize(hipStream_t s, int W, int K, int nchunk, int walker_type,
                            const double* Kl, const double* exx_part, double* E) {
  if (W <= 0) return;
  const double scl = (walker_type == AFQ_CLOSED) ? 2.0 : 1.0;
  hipLaunchKernelGGL(k_energy_finalize, dim3(W), dim3(64), 0, s, W, K, nchunk, -0.5 * scl,
                     0.5 * scl * scl, Kl, exx_part, E);
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
