// One-time construction of a system's derived tensors on the GPU:
//   haj[a][j]     = sum_i PsiT[a][i] h[i][j]               (half_rotate_hamiltonian, one_body.hpp:123-153)
//   Lnak[n][a][j] = sum_i PsiT[a][i] L[i][j][n]            (RealDenseHamiltonian.cpp:141-161)
//   v0[i][l]      = -1/2 sum_{j,n} L[i][j][n] L[l][j][n]   (RealDenseHamiltonian.cpp:163-180)
//   vMF           = Re vbias(G_trial)                      (NOMSD::vMF NOMSD.cpp:437-465,
//                                                           zero_imag AFQMCBasePropagator.cpp:107)
//   H1            = dt (h + v0) + vHS(vMF), hermitized     (getOneBodyPropagatorMatrix,
//                                                           Real3IndexFactorization.hpp:99-124)
//   P1            = exp(-H1/2)                             (generateP1, AFQMCBasePropagator.cpp:141-157)
// The reference takes exp_hermitian through an eigendecomposition on the host (exp.hpp:27-60);
// here P1 is the scaling-and-squaring Taylor exponential (degree 18 after scaling to norm <= 1/4,
// truncation < 1e-25 relative) on the GPU, which agrees with it to rounding.
#include "afq_internal.hpp"

#include <vector>

namespace afq {

namespace {

__global__ void k_half_rotate(int M, int N, int K, const cd* __restrict__ PsiT,
                              const double* __restrict__ h, const double* __restrict__ Likn,
                              cd* __restrict__ haj, cd* __restrict__ Lnak) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nL = (long)K * N * M;
  if (e < nL) {
    const int j = e % M;
    const int a = (e / M) % N;
    const int n = e / ((long)M * N);
    cd s = cmk(0, 0);
    for (int i = 0; i < M; ++i) s += Likn[((long)i * M + j) * K + n] * PsiT[(long)a * M + i];
    Lnak[e] = s;
  } else if (e < nL + (long)N * M) {
    const long f = e - nL;
    const int j = f % M, a = f / M;
    cd s = cmk(0, 0);
    for (int i = 0; i < M; ++i) s += h[(long)i * M + j] * PsiT[(long)a * M + i];
    haj[f] = s;
  }
}

__global__ void k_v0(int M, int K, const double* __restrict__ Likn, double* __restrict__ v0) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * M) return;
  const int i = e / M, l = e % M;
  double s = 0.0;
  for (int j = 0; j < M; ++j) {
    const double* a = Likn + ((long)i * M + j) * K;
    const double* b = Likn + ((long)l * M + j) * K;
    for (int n = 0; n < K; ++n) s += a[n] * b[n];
  }
  v0[e] = -0.5 * s;
}

__global__ void k_trial_phi(int M, int N, const cd* __restrict__ PsiT, cd* __restrict__ Phi0) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * N) return;
  const int i = e / N, a = e % N;
  Phi0[e] = cconj(PsiT[(long)a * M + i]);
}

__global__ void k_zero_imag(int n, cd* v) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) v[e].im = 0.0;
}

// H1 = dt (h + v0) + ms, hermitized (tensor.hpp:75-93); A = -H1/2
__global__ void k_build_H1(int M, double dt, const double* __restrict__ h,
                           const double* __restrict__ v0, const cd* __restrict__ ms,
                           cd* __restrict__ H1, cd* __restrict__ A) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * M) return;
  const int i = e / M, j = e % M;
  auto raw = [&](int r, int c) {
    const int q = r * M + c;
    return cmk(dt * h[q] + dt * v0[q], 0.0) + ms[q];
  };
  cd x;
  if (i == j) {
    x = cmk(raw(i, i).re, 0.0);
  } else if (i < j) {
    x = 0.5 * raw(i, j) + 0.5 * cconj(raw(j, i));
  } else {
    x = cconj(0.5 * raw(j, i) + 0.5 * cconj(raw(i, j)));
  }
  H1[e] = x;
  A[e] = -0.5 * x;
}

// C = alpha A B (+ I if add_identity); n x n, one thread per element
__global__ void k_zgemm_small(int n, double alpha, const cd* __restrict__ A,
                              const cd* __restrict__ B, int add_identity, cd* __restrict__ C) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int i = e / n, j = e % n;
  cd s = cmk(0, 0);
  for (int k = 0; k < n; ++k) s += A[(long)i * n + k] * B[(long)k * n + j];
  s = alpha * s;
  if (add_identity && i == j) s.re += 1.0;
  C[e] = s;
}

__global__ void k_to_f32(long n, const double* __restrict__ x, float* __restrict__ y) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = (float)x[e];
}

inline dim3 blocks(long n, int t = 256) { return dim3((unsigned)((n + t - 1) / t)); }

}  // namespace

void launch_half_rotate(hipStream_t s, int M, int N, int K, const double* PsiT, const double* h,
                        const double* Likn, double* haj, double* Lnak) {
  const long nL = (long)K * N * M;
  hipLaunchKernelGGL(k_half_rotate, blocks(nL + (long)N * M), dim3(256), 0, s, M, N, K,
                     (const cd*)PsiT, h, Likn, (cd*)haj, (cd*)Lnak);
  AFQ_HIP(hipGetLastError());
}

// v0[i][l] = -1/2 sum_{j,n} L[i][j][n] L[l][j][n]  (RealDenseHamiltonian.cpp:163-180)
void launch_v0(hipStream_t s, int M, int K, const double* Likn, double* v0) {
  hipLaunchKernelGGL(k_v0, blocks((long)M * M), dim3(256), 0, s, M, K, Likn, v0);
  AFQ_HIP(hipGetLastError());
}

// vMF <- Re vMF (zero_imag, AFQMCBasePropagator.cpp:107)
void launch_zero_imag(hipStream_t s, int n, double* v) {
  hipLaunchKernelGGL(k_zero_imag, blocks(n), dim3(256), 0, s, n, (cd*)v);
  AFQ_HIP(hipGetLastError());
}

void sparsify_one_body_propagator(int M, double* P) {
  const long n = (long)M * M;
  double vmax = 0.0, imax = 0.0;
  for (long e = 0; e < n; ++e) {
    const double a = std::hypot(P[2 * e], P[2 * e + 1]);
    if (a > 1e-8) {
      vmax = std::max(vmax, a);
      imax = std::max(imax, std::fabs(P[2 * e + 1]));
    } else {
      P[2 * e] = P[2 * e + 1] = 0.0;
    }
  }
  if (imax <= 1e-14 * vmax) {
    for (long e = 0; e < n; ++e) {
      P[2 * e + 1] = 0.0;
      if (!(std::fabs(P[2 * e]) > 1e-8)) P[2 * e] = 0.0;
    }
  }
}

// H1 = dt (h + v0) + vHS(vMF), hermitized, and P1 = exp(-H1/2) (getOneBodyPropagatorMatrix,
// Real3IndexFactorization.hpp:99-124; generateP1, AFQMCBasePropagator.cpp:141-157) by scaling
// and squaring of a degree-18 Taylor polynomial. Synchronises the stream (the scaling is chosen
// on the host from ||A||_1).
void build_one_body_propagator(hipStream_t s, int M, int K, double dt, const double* h,
                               const double* v0, const double* Likn, const double* vMF,
                               double* H1, double* P1) {
  // scratch: ms, A, P, Q [M*M] (c128)
  double* tmp = nullptr;
  AFQ_HIP(hipMalloc(&tmp, 2 * (size_t)M * M * 4 * sizeof(double)));
  double* ms = tmp;
  double* A = ms + 2 * (size_t)M * M;
  double* P = A + 2 * (size_t)M * M;
  double* Q = P + 2 * (size_t)M * M;
  try {
    launch_gemm_nt_z(s, AFQ_FP64, 1, M * M, K, std::sqrt(dt), vMF, K, Likn, K, false, ms,
                     (long)M * M);
    hipLaunchKernelGGL(k_build_H1, blocks((long)M * M), dim3(256), 0, s, M, dt, h, v0,
                       (const cd*)ms, (cd*)H1, (cd*)A);
    AFQ_HIP(hipGetLastError());
    // ||A||_1 on the host chooses the scaling
    std::vector<double> hA(2 * (size_t)M * M);
    AFQ_HIP(hipMemcpyAsync(hA.data(), A, hA.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    AFQ_HIP(hipStreamSynchronize(s));
    double nrm = 0.0;
    for (int j = 0; j < M; ++j) {
      double c = 0.0;
      for (int i = 0; i < M; ++i) c += std::hypot(hA[2 * ((size_t)i * M + j)], hA[2 * ((size_t)i * M + j) + 1]);
      nrm = std::max(nrm, c);
    }
    int sq = 0;
    while (nrm > 0.25) {
      nrm *= 0.5;
      ++sq;
    }
    const double bscale = std::ldexp(1.0, -sq);
    // Horner: P = I + B P / k for k = 18..1, B = A / 2^sq
    std::vector<double> eye(2 * (size_t)M * M, 0.0);
    for (int i = 0; i < M; ++i) eye[2 * ((size_t)i * M + i)] = 1.0;
    AFQ_HIP(hipMemcpyAsync(P, eye.data(), eye.size() * sizeof(double), hipMemcpyHostToDevice, s));
    AFQ_HIP(hipStreamSynchronize(s));
    for (int k = 18; k >= 1; --k) {
      hipLaunchKernelGGL(k_zgemm_small, blocks((long)M * M), dim3(256), 0, s, M, bscale / k,
                         (const cd*)A, (const cd*)P, 1, (cd*)Q);
      std::swap(P, Q);
    }
    for (int q = 0; q < sq; ++q) {
      hipLaunchKernelGGL(k_zgemm_small, blocks((long)M * M), dim3(256), 0, s, M, 1.0,
                         (const cd*)P, (const cd*)P, 0, (cd*)Q);
      std::swap(P, Q);
    }
    // the reference keeps P1 sparse (denseP1 = false, parameters.hpp:120): entries with
    // |x| <= 1e-8 are dropped (to_csr, AFQMCBasePropagator.cpp:165-166), and a P1 whose
    // imaginary parts are all within 1e-14 of its largest entry is applied in real arithmetic
    // (AFQMCBasePropagator.cpp:174-211): the same matrix here, with those entries zeroed
    std::vector<double> hP(2 * (size_t)M * M);
    AFQ_HIP(hipMemcpyAsync(hP.data(), P, hP.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    AFQ_HIP(hipStreamSynchronize(s));
    sparsify_one_body_propagator(M, hP.data());
    AFQ_HIP(hipMemcpyAsync(P1, hP.data(), hP.size() * sizeof(double), hipMemcpyHostToDevice, s));
    AFQ_HIP(hipGetLastError());
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(tmp);
    throw;
  }
  AFQ_HIP(hipFree(tmp));
}

void setup_hamiltonian(afq_ham_s* H, hipStream_t s) {
  const int M = H->M, N = H->N, K = H->K;
  const long nL = (long)K * N * M;
  hipLaunchKernelGGL(k_half_rotate, blocks(nL + (long)N * M), dim3(256), 0, s, M, N, K,
                     (const cd*)H->PsiT, H->h, H->Likn, (cd*)H->haj, (cd*)H->Lnak);
  launch_v0(s, M, K, H->Likn, H->v0);
  AFQ_HIP(hipGetLastError());

  // scratch: Phi0 [M*N], Gt [N*M] (c128)
  double* tmp = nullptr;
  AFQ_HIP(hipMalloc(&tmp, 2 * (size_t)M * N * 2 * sizeof(double)));
  double* Phi0 = tmp;
  double* Gt = Phi0 + 2 * (size_t)M * N;
  try {
    hipLaunchKernelGGL(k_trial_phi, blocks((long)M * N), dim3(256), 0, s, M, N,
                       (const cd*)H->PsiT, (cd*)Phi0);
    launch_green(s, 1, M, N, H->PsiT, Phi0, (long)M * N, Gt, nullptr, 0, 1.0, cmk(0, 0),
                 nullptr, 0.0, 0.0, nullptr, 0);
    const double vscale = std::sqrt(H->dt) * (H->walker_type == AFQ_CLOSED ? 2.0 : 1.0);
    launch_gemm_nt_z(s, AFQ_FP64, 1, K, N * M, vscale, Gt, (long)N * M, H->Lnak, (long)N * M,
                     true, H->vMF, K);
    launch_zero_imag(s, K, H->vMF);
    build_one_body_propagator(s, M, K, H->dt, H->h, H->v0, H->Likn, H->vMF, H->H1, H->P1);
    // FP32 copies of the Cholesky tensors for the FP32 contraction form
    hipLaunchKernelGGL(k_to_f32, blocks(2 * nL), dim3(256), 0, s, 2 * nL, H->Lnak, H->Lnak_f);
    hipLaunchKernelGGL(k_to_f32, blocks((long)M * M * K), dim3(256), 0, s, (long)M * M * K,
                       H->Likn, H->Likn_f);
    AFQ_HIP(hipGetLastError());
    AFQ_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(tmp);
    throw;
  }
  AFQ_HIP(hipFree(tmp));
}

}  // namespace afq
