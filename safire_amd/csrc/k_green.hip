// Batched mixed density matrix and log-overlap of closed-shell walkers (one workgroup per walker).
//
// det_ops::MixedDensityMatrix / detail::log_overlap_impl (density_matrix.hpp:113-156, 535-601):
//   T = PsiT Phi_w               [N,N]
//   LU with partial pivoting     (LAPACK getrf: pivot = first max of |re|+|im|)
//   ov += sum_i log(+-U_ii + DBL_MIN), imaginary part wrapped to [-pi,pi]
//                                (log_determinant_from_getrf, determinants_impl.hpp:32-75)
//   T^-1, all zero when some U_ii == 0   (getri_or_zero, nda_functions.hpp:422-438)
//   G = T^-T Phi_w^T             [N,M] compact Green's function
// Optional epilogue: E1 = E0 + scl * sum_{a,j} G[a][j] haj[a][j] (Real3IndexFactorization.hpp:505-508).
//
// The walker's Slater matrix, T and T^-1 live in LDS; the LU runs with one barrier per
// elimination step. Small N (6..32) makes this kernel latency-bound; it carries ~3% of a step's
// flops.
#include "afq_internal.hpp"

#include <cfloat>

namespace afq {

namespace {

constexpr int kThreads = 256;

__device__ inline double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

__global__ __launch_bounds__(kThreads) void k_green(int M, int N, const cd* __restrict__ PsiT,
                                                    const cd* __restrict__ Phi, long ldw,
                                                    cd* __restrict__ G, cd* __restrict__ ov,
                                                    int ov_mode, double ov_scale, cd ov_add,
                                                    const cd* __restrict__ haj, double E0,
                                                    double e1_scale, cd* __restrict__ E,
                                                    long ldE) {
  extern __shared__ cd smem[];
  cd* sPhi = smem;           // [M][N]
  cd* sT = sPhi + M * N;     // [N][N]  LU factors
  cd* sX = sT + N * N;       // [N][N]  inverse
  __shared__ int ipiv[64];
  __shared__ int singular;
  __shared__ double red[2 * (kThreads / 64)];

  const int w = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const cd* P = Phi + (long)w * ldw;

  for (int e = tid; e < M * N; e += kThreads) sPhi[e] = P[e];
  __syncthreads();

  // T[a][b] = sum_i PsiT[a][i] Phi[i][b]
  for (int e = tid; e < N * N; e += kThreads) {
    const int a = e / N, b = e - a * N;
    const cd* pa = PsiT + (long)a * M;
    cd s = cmk(0.0, 0.0);
    for (int i = 0; i < M; ++i) s += pa[i] * sPhi[i * N + b];
    sT[e] = s;
  }
  __syncthreads();

  // right-looking LU with partial pivoting (zgetf2 order of operations)
  for (int j = 0; j < N; ++j) {
    if (tid < 64) {
      const int i = j + lane;
      double v = (i < N) ? cabs1(sT[i * N + j]) : -1.0;
      int idx = (i < N) ? i : 0x7fffffff;
      for (int off = 32; off > 0; off >>= 1) {
        double ov_ = shfl_xor_d(v, off);
        int oi = __shfl_xor(idx, off, 64);
        if (ov_ > v || (ov_ == v && oi < idx)) {
          v = ov_;
          idx = oi;
        }
      }
      if (lane == 0) ipiv[j] = idx;
    }
    __syncthreads();
    const int p = ipiv[j];
    if (p != j)
      for (int c = tid; c < N; c += kThreads) {
        cd t = sT[j * N + c];
        sT[j * N + c] = sT[p * N + c];
        sT[p * N + c] = t;
      }
    __syncthreads();
    const cd piv = sT[j * N + j];
    if (!cis0(piv)) {
      const cd r = cdiv(cmk(1.0, 0.0), piv);
      for (int i = j + 1 + tid; i < N; i += kThreads) sT[i * N + j] = sT[i * N + j] * r;
    }
    __syncthreads();
    const int nt = N - 1 - j;
    for (int e = tid; e < nt * nt; e += kThreads) {
      const int i = j + 1 + e / nt, c = j + 1 + e % nt;
      const cd l = sT[i * N + j];
      if (!cis0(l)) sT[i * N + c] -= l * sT[j * N + c];
    }
    __syncthreads();
  }

  if (tid == 0) {
    cd acc = (ov_mode == 2) ? ov[w] : cmk(0.0, 0.0);
    int sing = 0;
    for (int i = 0; i < N; ++i) {
      const cd u = sT[i * N + i];
      if (cis0(u)) sing = 1;
      const cd x = (ipiv[i] == i) ? u : -u;
      acc += clog_(cmk(x.re + DBL_MIN, x.im));
    }
    const double pi = 3.14159265358979323846;
    // the reference loops until the value is in range; a non-finite accumulated input would never
    // get there, so only finite values are wrapped
    if (!isfinite(acc.im)) {
    } else if (acc.im > pi) {
      while (acc.im > pi) acc.im -= 2 * pi;
    } else if (acc.im < -pi) {
      while (acc.im < -pi) acc.im += 2 * pi;
    }
    if (ov_mode != 0) ov[w] = ov_scale * acc + ov_add;
    singular = sing;
  }
  __syncthreads();
  if (G == nullptr && E == nullptr) return;

  // T^-1 by column: X = U^-1 L^-1 P e_c  (the getri_or_zero result)
  if (!singular && tid < N) {
    const int c = tid;
    for (int i = 0; i < N; ++i) sX[i * N + c] = cmk(i == c ? 1.0 : 0.0, 0.0);
    for (int i = 0; i < N; ++i) {
      const int p = ipiv[i];
      if (p != i) {
        cd t = sX[i * N + c];
        sX[i * N + c] = sX[p * N + c];
        sX[p * N + c] = t;
      }
    }
    for (int i = 0; i < N; ++i) {
      cd s = sX[i * N + c];
      for (int k = 0; k < i; ++k) s -= sT[i * N + k] * sX[k * N + c];
      sX[i * N + c] = s;
    }
    for (int i = N - 1; i >= 0; --i) {
      cd s = sX[i * N + c];
      for (int k = i + 1; k < N; ++k) s -= sT[i * N + k] * sX[k * N + c];
      sX[i * N + c] = cdiv(s, sT[i * N + i]);
    }
  }
  __syncthreads();

  // G[a][i] = sum_b X[b][a] Phi[i][b];  E1 partial sum_{a,i} G[a][i] haj[a][i]
  cd e1 = cmk(0.0, 0.0);
  cd* Gw = G ? G + (long)w * N * M : nullptr;
  for (int e = tid; e < N * M; e += kThreads) {
    const int a = e / M, i = e - a * M;
    cd s = cmk(0.0, 0.0);
    if (!singular)
      for (int b = 0; b < N; ++b) s += sX[b * N + a] * sPhi[i * N + b];
    if (Gw) Gw[e] = s;
    if (haj) e1 += s * haj[e];
  }
  if (E != nullptr && haj != nullptr) {
    for (int off = 32; off > 0; off >>= 1) {
      e1.re += shfl_xor_d(e1.re, off);
      e1.im += shfl_xor_d(e1.im, off);
    }
    if (lane == 0) {
      red[2 * (tid >> 6)] = e1.re;
      red[2 * (tid >> 6) + 1] = e1.im;
    }
    __syncthreads();
    if (tid == 0) {
      cd s = cmk(0.0, 0.0);
      for (int q = 0; q < kThreads / 64; ++q) s += cmk(red[2 * q], red[2 * q + 1]);
      E[(long)w * ldE] = cmk(E0, 0.0) + e1_scale * s;
    }
  }
}

}  // namespace

void launch_green(hipStream_t s, int nb, int M, int N, const double* PsiT, const double* Phi,
                  long ldw, double* G, double* ov, int ov_mode, double ov_scale, cd ov_add,
                  const double* haj, double E0, double e1_scale, double* E, long ldE) {
  if (nb <= 0 || N <= 0) return;
  AFQ_CHECK(N <= 64, AFQ_EUNSUPPORTED, "mixed density matrix: N > 64 electrons per spin");
  const size_t lds = sizeof(cd) * ((size_t)M * N + 2 * (size_t)N * N);
  AFQ_CHECK(lds <= 160 * 1024 - 1024, AFQ_EUNSUPPORTED,
            "mixed density matrix: walker does not fit in LDS");
  if (lds > 64 * 1024)
    AFQ_HIP(hipFuncSetAttribute((const void*)k_green,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_green, dim3(nb), dim3(kThreads), lds, s, M, N,
                     reinterpret_cast<const cd*>(PsiT), reinterpret_cast<const cd*>(Phi), ldw,
                     reinterpret_cast<cd*>(G), reinterpret_cast<cd*>(ov), ov_mode, ov_scale,
                     ov_add, reinterpret_cast<const cd*>(haj), E0, e1_scale,
                     reinterpret_cast<cd*>(E), ldE);
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
