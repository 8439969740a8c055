// Batched re-orthogonalisation of walkers (det_ops::orthogonalize, orthogonalize.hpp:131-169):
// Householder QR of Phi_w [M,N] with LAPACK zgeqrf conventions (zlarfg: real
// beta = -sign(Re alpha) |x|, tau = (beta - alpha)/beta, v = x/(alpha - beta)), log det R
// accumulated as log_determinant_from_geqrf does (determinants_impl.hpp:77-108), Q formed by
// backward accumulation (zungqr), and Q's columns scaled by sign(Re R_ii): the result is the
// unique Q whose R has a positive real diagonal.
//
// One workgroup per walker; A (the reflectors) and Q live in LDS with rows padded to N + 1
// complex numbers, so that the lanes of a column access (consecutive rows) fall into different
// banks instead of all into one. Each elimination step reduces a column norm in one wavefront and
// applies the reflector with one wavefront per trailing column (lanes over rows, shuffle
// reductions).
#include "afq_internal.hpp"

#include <cfloat>

namespace afq {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ inline cd wave_sum(cd v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.re += __shfl_xor(v.re, off, 64);
    v.im += __shfl_xor(v.im, off, 64);
  }
  return v;
}

__global__ __launch_bounds__(kThreads) void k_qr(int M, int N, cd* __restrict__ Phi, long ldw,
                                                 cd* __restrict__ logdetR) {
  extern __shared__ cd smem[];
  const int ld = N + 1;    // padded row stride
  cd* A = smem;            // [M][ld]
  cd* Q = A + M * ld;      // [M][ld]
  __shared__ cd tau[64];
  __shared__ double scl[64];
  const int w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cd* P = Phi + (long)w * ldw;
  for (int e = tid; e < M * N; e += kThreads) A[(e / N) * ld + e % N] = P[e];
  __syncthreads();
  const int kmax = min(M, N);

  for (int j = 0; j < kmax; ++j) {
    if (wave == 0) {
      double x2 = 0.0;
      for (int i = j + 1 + lane; i < M; i += 64) x2 += cnorm(A[i * ld + j]);
      for (int off = 32; off > 0; off >>= 1) x2 += __shfl_xor(x2, off, 64);
      const cd alpha = A[j * ld + j];
      cd t = cmk(0.0, 0.0), sc = cmk(0.0, 0.0);
      double beta = 0.0;
      const bool identity = (x2 == 0.0 && alpha.im == 0.0);
      if (!identity) {
        const double nrm = sqrt(cnorm(alpha) + x2);
        beta = (alpha.re >= 0.0) ? -nrm : nrm;
        t = cmk((beta - alpha.re) / beta, -alpha.im / beta);
        sc = cdiv(cmk(1.0, 0.0), cmk(alpha.re - beta, alpha.im));
      }
      for (int i = j + 1 + lane; i < M; i += 64)
        if (!identity) A[i * ld + j] = A[i * ld + j] * sc;
      if (lane == 0) {
        tau[j] = t;
        if (!identity) A[j * ld + j] = cmk(beta, 0.0);
      }
    }
    __syncthreads();
    // A[j:, c] -= conj(tau) v (v^H A[j:, c]) for c > j
    const cd tj = cconj(tau[j]);
    for (int c = j + 1 + wave; c < N; c += kWaves) {
      cd s = cmk(0.0, 0.0);
      for (int i = j + 1 + lane; i < M; i += 64) s += cconj(A[i * ld + j]) * A[i * ld + c];
      s = wave_sum(s);
      s = (A[j * ld + c] + s);
      s = tj * s;
      if (lane == 0) A[j * ld + c] -= s;
      for (int i = j + 1 + lane; i < M; i += 64) A[i * ld + c] -= s * A[i * ld + j];
    }
    __syncthreads();
  }

  if (tid < kmax) scl[tid] = (A[tid * ld + tid].re < 0.0) ? -1.0 : 1.0;
  if (logdetR != nullptr && tid == 0) {
    cd acc = logdetR[w];
    for (int i = 0; i < kmax; ++i) {
      const cd r = (A[i * ld + i].re < 0.0) ? -A[i * ld + i] : A[i * ld + i];
      acc += clog_(cmk(r.re + DBL_MIN, r.im));
    }
    logdetR[w] = acc;
  }
  for (int e = tid; e < M * N; e += kThreads) {
    const int i = e / N, c = e - i * N;
    Q[i * ld + c] = cmk(i == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int j = kmax - 1; j >= 0; --j) {
    const cd t = tau[j];
    for (int c = j + wave; c < N; c += kWaves) {
      cd s = cmk(0.0, 0.0);
      for (int i = j + 1 + lane; i < M; i += 64) s += cconj(A[i * ld + j]) * Q[i * ld + c];
      s = wave_sum(s);
      s = t * (Q[j * ld + c] + s);
      if (lane == 0) Q[j * ld + c] -= s;
      for (int i = j + 1 + lane; i < M; i += 64) Q[i * ld + c] -= s * A[i * ld + j];
    }
    __syncthreads();
  }
  for (int e = tid; e < M * N; e += kThreads) {
    const int i = e / N, c = e % N;
    P[e] = (c < kmax) ? scl[c] * Q[i * ld + c] : Q[i * ld + c];
  }
}

}  // namespace

void launch_orthogonalize(hipStream_t s, int nb, int M, int N, double* Phi, long ldw,
                          double* logdetR) {
  if (nb <= 0 || (long)M * N == 0) return;
  AFQ_CHECK(N <= 64, AFQ_EUNSUPPORTED, "orthogonalize: N > 64");
  const size_t lds = 2 * sizeof(cd) * (size_t)M * (N + 1);
  AFQ_CHECK(lds <= 160 * 1024 - 2048, AFQ_EUNSUPPORTED, "orthogonalize: walker too large");
  ensure_dynamic_lds((const void*)k_qr, lds);
  hipLaunchKernelGGL(k_qr, dim3(nb), dim3(kThreads), lds, s, M, N, reinterpret_cast<cd*>(Phi),
                     ldw, reinterpret_cast<cd*>(logdetR));
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
