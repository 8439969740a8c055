// afq_oracle.cpp -- FP64 CPU restatement of SAFIRE's HOST_MEMORY path for the phaseless-AFQMC
// walker step with a Cholesky-factorised (Real3Index) Hamiltonian.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product library (safire_amd/csrc) never does.
//
// Every function cites the reference file:line it restates. Layouts are the reference's: C order,
// complex<double> interleaved, walker matrices Phi[w][i][a] (i orbital, a electron) with a walker
// stride ldw (the walker_buffer row), compact G[w][a][k], vHS V[w][i][j], E[w][3] = (E1,EXX,EJ).
//
// Pinning (see tests/test_oracle_*.py): getrf/logdet/getri/MixedDensityMatrix/Log_Overlap against
// the literal known-answer values of tests/test_sdet_ops.cpp; determinants against permutation
// expansion; vbias/energy against Wick's theorem on the explicit two-electron integrals and the
// closed-form Hubbard Cholesky vectors; propagate against scipy's expm; orthogonalize against
// numpy's QR and the Log_Overlap(Q,Q)=0 invariant; exp_hermitian against scipy's expm.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>
#include <numeric>
#include <omp.h>

typedef std::complex<double> cd;

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
inline double cabs1(cd z) { return std::fabs(z.real()) + std::fabs(z.imag()); }

// ---------------------------------------------------------------------------------------------
// Dense LU with partial pivoting, LAPACK zgetf2 semantics (pivot = first max of |re|+|im|,
// izamax); ipiv is 1-based as LAPACK returns it. Returns info (0, or j+1 for the first exactly
// zero pivot U(j,j)). Row-major n x n, in place.
// Reference use: nda::lapack::getrf in density_matrix.hpp:68,103,147.
int getrf(int n, cd* A, int* ipiv) {
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = cabs1(A[j * n + j]);
    for (int i = j + 1; i < n; ++i) {
      double v = cabs1(A[i * n + j]);
      if (v > best) { best = v; p = i; }
    }
    ipiv[j] = p + 1;
    if (A[p * n + j] != cd(0.0)) {
      if (p != j)
        for (int c = 0; c < n; ++c) std::swap(A[j * n + c], A[p * n + c]);
      cd r = 1.0 / A[j * n + j];
      for (int i = j + 1; i < n; ++i) A[i * n + j] *= r;
    } else if (info == 0) {
      info = j + 1;
    }
    for (int i = j + 1; i < n; ++i) {
      cd l = A[i * n + j];
      if (l == cd(0.0)) continue;
      for (int c = j + 1; c < n; ++c) A[i * n + c] -= l * A[j * n + c];
    }
  }
  return info;
}

// log_determinant_from_getrf_impl, determinants_impl.hpp:32-75: res += sum_i log(+-U_ii + DBL_MIN)
// with the sign flipped where row i was interchanged, then the imaginary part of the accumulated
// value is brought to [-pi, pi].
void log_det_from_getrf(int n, const cd* LU, const int* ipiv, cd& res) {
  const cd small(std::numeric_limits<double>::min(), 0.0);
  for (int i = 0; i < n; ++i) {
    cd u = LU[i * n + i];
    if (ipiv[i] == i + 1) res += std::log(u + small);
    else                  res += std::log(-u + small);
  }
  if (res.imag() > kPi) {
    while (res.imag() > kPi) res.imag(res.imag() - 2 * kPi);
  } else if (res.imag() < -kPi) {
    while (res.imag() < -kPi) res.imag(res.imag() + 2 * kPi);
  }
}

// getri_or_zero, nda_functions.hpp:422-438: inverse from the LU factors; a singular matrix
// (an exactly zero U_ii, where LAPACK getri reports info>0) gives an all-zero result.
int getri_or_zero(int n, cd* LU, const int* ipiv) {
  for (int i = 0; i < n; ++i)
    if (LU[i * n + i] == cd(0.0)) {
      std::fill(LU, LU + (size_t)n * n, cd(0.0));
      return i + 1;
    }
  // Solve A X = I with A = P L U (rows permuted by ipiv in order).
  std::vector<cd> X((size_t)n * n, cd(0.0));
  for (int i = 0; i < n; ++i) X[i * n + i] = 1.0;
  for (int i = 0; i < n; ++i) {
    int p = ipiv[i] - 1;
    if (p != i)
      for (int c = 0; c < n; ++c) std::swap(X[i * n + c], X[p * n + c]);
  }
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) {  // L y = b (unit lower)
      cd s = X[i * n + c];
      for (int k = 0; k < i; ++k) s -= LU[i * n + k] * X[k * n + c];
      X[i * n + c] = s;
    }
    for (int i = n - 1; i >= 0; --i) {  // U x = y
      cd s = X[i * n + c];
      for (int k = i + 1; k < n; ++k) s -= LU[i * n + k] * X[k * n + c];
      X[i * n + c] = s / LU[i * n + i];
    }
  }
  std::copy(X.begin(), X.end(), LU);
  return 0;
}

// Complex GEMM helpers with explicit real arithmetic (row-major).
// C[m][n] = alpha * sum_k A[m][k] * B[k][n] + beta*C
void zgemm_nn(int m, int n, int k, cd alpha, const cd* A, long lda, const cd* B, long ldb,
              cd beta, cd* C, long ldc) {
  for (int i = 0; i < m; ++i) {
    cd* c = C + (long)i * ldc;
    if (beta == cd(0.0)) std::fill(c, c + n, cd(0.0));
    else if (beta != cd(1.0)) for (int j = 0; j < n; ++j) c[j] *= beta;
    double* cr = reinterpret_cast<double*>(c);
    for (int p = 0; p < k; ++p) {
      cd a = alpha * A[(long)i * lda + p];
      const double ar = a.real(), ai = a.imag();
      const double* b = reinterpret_cast<const double*>(B + (long)p * ldb);
      for (int j = 0; j < n; ++j) {
        double br = b[2 * j], bi = b[2 * j + 1];
        cr[2 * j] += ar * br - ai * bi;
        cr[2 * j + 1] += ar * bi + ai * br;
      }
    }
  }
}

// C[m][n] = sum_k A[m][k] * B[n][k]   (A times transpose(B), no conjugation)
void zgemm_nt(int m, int n, int k, cd alpha, const cd* A, long lda, const cd* B, long ldb,
              cd* C, long ldc) {
  for (int i = 0; i < m; ++i) {
    const double* a = reinterpret_cast<const double*>(A + (long)i * lda);
    for (int j = 0; j < n; ++j) {
      const double* b = reinterpret_cast<const double*>(B + (long)j * ldb);
      double sr = 0, si = 0;
      for (int p = 0; p < k; ++p) {
        double ar = a[2 * p], ai = a[2 * p + 1], br = b[2 * p], bi = b[2 * p + 1];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      C[(long)i * ldc + j] = alpha * cd(sr, si);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Real symmetric cyclic Jacobi eigensolver (Golub & Van Loan, Alg. 8.5.3). A is n x n
// row-major, destroyed; on return w holds eigenvalues and V (row-major) the eigenvectors in
// columns.
void jacobi_eigh(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double v = A[i * n + j] * A[i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (apq == 0.0) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double tau = (aqq - app) / (2.0 * apq);
        double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

}  // namespace

extern "C" {

// ============================= det_ops (SlaterDeterminantOperations) =========================

int orc_getrf(int n, cd* A, int* ipiv) { return getrf(n, A, ipiv); }

void orc_log_det_from_getrf(int n, const cd* LU, const int* ipiv, cd* res) {
  log_det_from_getrf(n, LU, ipiv, *res);
}

int orc_getri_or_zero(int n, cd* LU, const int* ipiv) { return getri_or_zero(n, LU, ipiv); }

// det_ops::detail::inverse_logdet (density_matrix.hpp:80-113), batched: logdet accumulates.
void orc_inverse_logdet(int nb, int n, const cd* A, cd* Ainv, cd* logdet) {
  std::vector<int> ipiv(n);
  for (int b = 0; b < nb; ++b) {
    cd* T = Ainv + (size_t)b * n * n;
    std::copy(A + (size_t)b * n * n, A + (size_t)(b + 1) * n * n, T);
    getrf(n, T, ipiv.data());
    log_det_from_getrf(n, T, ipiv.data(), logdet[b]);
    getri_or_zero(n, T, ipiv.data());
  }
}

// detail::log_overlap_impl (density_matrix.hpp:113-156, herm=true) for one walker:
// T = PsiT * Phi  ([N,M] x [M,N]), LU, ov += wrapped log det; optionally T <- T^{-1}.
static void log_overlap_one(int M, int N, const cd* PsiT, const cd* Phi, cd* T, int* ipiv,
                            cd& ov, bool invert) {
  zgemm_nn(N, N, M, 1.0, PsiT, M, Phi, N, 0.0, T, N);
  getrf(N, T, ipiv);
  log_det_from_getrf(N, T, ipiv, ov);
  if (invert) getri_or_zero(N, T, ipiv);
}

// det_ops::Log_Overlap (density_matrix.hpp:437-447): ov[w] += log det(PsiT * Phi_w).
void orc_log_overlap(int nb, int M, int N, const cd* PsiT, const cd* Phi, long ldw, cd* ov) {
  if (N == 0) return;
#pragma omp parallel
  {
    std::vector<cd> T((size_t)N * N);
    std::vector<int> ipiv(N);
#pragma omp for schedule(static)
    for (int w = 0; w < nb; ++w)
      log_overlap_one(M, N, PsiT, Phi + (long)w * ldw, T.data(), ipiv.data(), ov[w], false);
  }
}

// det_ops::MixedDensityMatrix (density_matrix.hpp:535-601, herm=true).
// compact: G[w] = T^{-T} Phi_w^T  -> [N, M]; full: G[w] = PsiT^T (T^{-T} Phi^T) -> [M, M].
void orc_mixed_density_matrix(int nb, int M, int N, const cd* PsiT, const cd* Phi, long ldw,
                              cd* G, cd* ov, int compact) {
  const long gsz = compact ? (long)N * M : (long)M * M;
  if (N == 0) {
    if (!compact) std::fill(G, G + nb * gsz, cd(0.0));
    return;
  }
#pragma omp parallel
  {
    std::vector<cd> T((size_t)N * N), TNM((size_t)N * M);
    std::vector<int> ipiv(N);
#pragma omp for schedule(static)
    for (int w = 0; w < nb; ++w) {
      const cd* P = Phi + (long)w * ldw;
      log_overlap_one(M, N, PsiT, P, T.data(), ipiv.data(), ov[w], true);
      // TNM[a][i] = sum_b Tinv[b][a] * Phi[i][b]
      cd* out = compact ? G + w * gsz : TNM.data();
      for (int a = 0; a < N; ++a)
        for (int i = 0; i < M; ++i) {
          cd s = 0.0;
          for (int b = 0; b < N; ++b) s += T[b * N + a] * P[(long)i * N + b];
          out[(long)a * M + i] = s;
        }
      if (!compact) {
        // C = transpose(PsiT) * TNM : C[i][j] = sum_a PsiT[a][i] * TNM[a][j]
        cd* C = G + w * gsz;
        for (int i = 0; i < M; ++i)
          for (int j = 0; j < M; ++j) {
            cd s = 0.0;
            for (int a = 0; a < N; ++a) s += PsiT[(long)a * M + i] * TNM[(long)a * M + j];
            C[(long)i * M + j] = s;
          }
      }
    }
  }
}

// ============================= Hamiltonian setup ============================================
// RealDenseHamiltonian::getHamiltonianOperations (RealDenseHamiltonian.cpp:135-176) and
// half_rotate_hamiltonian (one_body.hpp:123-153), one spin sector, spin-independent L:
//   haj[a][j]    = sum_i PsiT[a][i] h[i][j]
//   Lnak[n][a][j]= sum_i PsiT[a][i] L[i][j][n]
//   v0[i][l]     = -1/2 sum_{j,n} L[i][j][n] L[l][j][n]
void orc_half_rotate(int M, int N, int K, const double* h, const double* Likn, const cd* PsiT,
                     cd* haj, cd* Lnak) {
  for (int a = 0; a < N; ++a)
    for (int j = 0; j < M; ++j) {
      cd s = 0.0;
      for (int i = 0; i < M; ++i) s += PsiT[(long)a * M + i] * h[(long)i * M + j];
      haj[(long)a * M + j] = s;
    }
#pragma omp parallel for schedule(static)
  for (int n = 0; n < K; ++n)
    for (int a = 0; a < N; ++a)
      for (int j = 0; j < M; ++j) {
        cd s = 0.0;
        for (int i = 0; i < M; ++i) s += PsiT[(long)a * M + i] * Likn[((long)i * M + j) * K + n];
        Lnak[((long)n * N + a) * M + j] = s;
      }
}

void orc_exchange_v0(int M, int K, const double* Likn, double* v0) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; ++i)
    for (int l = 0; l < M; ++l) {
      double s = 0.0;
      for (int j = 0; j < M; ++j)
        for (int n = 0; n < K; ++n)
          s += Likn[((long)i * M + j) * K + n] * Likn[((long)l * M + j) * K + n];
      v0[(long)i * M + l] = -0.5 * s;
    }
}

// ============================= HamiltonianOperations (Real3IndexFactorization) ================
// vbias (Real3IndexFactorization.hpp:328-361, ndet==1 branch): v[w][n] = a sum_{a,k} G[w][a][k]
// Lnak[n][a][k] with a = sqrt(dt) * (CLOSED ? 2 : 1). nel rows of G per walker; for COLLINEAR
// the G rows of the down spin (rows nup..nup+ndn) contract with Lnak_dn.
void orc_vbias(int W, int M, int nup, int ndn, int K, int walker_type, double dt,
               const cd* Lnak_up, const cd* Lnak_dn, const cd* G, cd* v) {
  const int nel = (walker_type == 2) ? nup + ndn : nup;
  const double a = std::sqrt(dt) * (walker_type == 1 ? 2.0 : 1.0);
#pragma omp parallel for schedule(static)
  for (int w = 0; w < W; ++w) {
    const cd* Gw = G + (long)w * nel * M;
    for (int n = 0; n < K; ++n) {
      cd s = 0.0;
      const cd* L = Lnak_up + (long)n * nup * M;
      for (long e = 0; e < (long)nup * M; ++e) s += Gw[e] * L[e];
      if (walker_type == 2 && ndn > 0) {
        const cd* Ld = Lnak_dn + (long)n * ndn * M;
        for (long e = 0; e < (long)ndn * M; ++e) s += Gw[(long)nup * M + e] * Ld[e];
      }
      v[(long)w * K + n] = a * s;
    }
  }
}

// vHS (Real3IndexFactorization.hpp:282-313, host branch): V[w][i][j] = sqrt(dt) sum_n
// L[i][j][n] X[w][n].
void orc_vhs(int W, int M, int K, double dt, const double* Likn, const cd* X, cd* V) {
  const double a = std::sqrt(dt);
#pragma omp parallel for schedule(static)
  for (int w = 0; w < W; ++w) {
    const double* x = reinterpret_cast<const double*>(X + (long)w * K);
    for (long ij = 0; ij < (long)M * M; ++ij) {
      const double* L = Likn + ij * K;
      double sr = 0, si = 0;
      for (int n = 0; n < K; ++n) { sr += L[n] * x[2 * n]; si += L[n] * x[2 * n + 1]; }
      V[(long)w * M * M + ij] = cd(a * sr, a * si);
    }
  }
}

// energy / energy_impl (Real3IndexFactorization.hpp:138-174, 476-559, host branch):
//   E1  = E0 + scl sum_{a,j} G[a][j] haj[a][j]
//   T[b][n][a] = sum_k G[b][k] Lnak[n][a][k]        (scl = 2 CLOSED, 1 otherwise)
//   EXX = -scl/2 sum_{a,b,n} T[a][n][b] T[b][n][a]  (non-conjugated, nda::blas::dot)
//   Kl[n] = sum_a T[a][n][a] (summed over spins);  EJ = (CLOSED ? 2 : 1/2) sum_n Kl[n]^2
void orc_energy(int W, int M, int nup, int ndn, int K, int walker_type, double E0,
                const cd* haj_up, const cd* haj_dn, const cd* Lnak_up, const cd* Lnak_dn,
                const cd* G, cd* E) {
  const int nel = (walker_type == 2) ? nup + ndn : nup;
  const double scl = (walker_type == 1) ? 2.0 : 1.0;
  const int nspin = (walker_type == 2) ? 2 : 1;
#pragma omp parallel
  {
    std::vector<cd> T, Kl(K);
#pragma omp for schedule(static)
    for (int w = 0; w < W; ++w) {
      cd e1 = E0, exx = 0.0;
      std::fill(Kl.begin(), Kl.end(), cd(0.0));
      for (int is = 0; is < nspin; ++is) {
        const int ne = is == 0 ? nup : ndn;
        if (ne == 0) continue;
        const cd* Gs = G + (long)w * nel * M + (is == 0 ? 0 : (long)nup * M);
        const cd* haj = is == 0 ? haj_up : haj_dn;
        const cd* Lnak = is == 0 ? Lnak_up : Lnak_dn;
        cd s1 = 0.0;
        for (long e = 0; e < (long)ne * M; ++e) s1 += Gs[e] * haj[e];
        e1 += scl * s1;
        // T2d[b][(n,a)] = G[b][:] . Lnak[(n,a)][:]
        T.resize((size_t)ne * K * ne);
        zgemm_nt(ne, K * ne, M, 1.0, Gs, M, Lnak, M, T.data(), (long)K * ne);
        auto Tv = [&](int b, int n, int a) { return T[((size_t)b * K + n) * ne + a]; };
        for (int ia = 0; ia < ne; ++ia) {
          cd d = 0.0;
          for (int n = 0; n < K; ++n) d += Tv(ia, n, ia) * Tv(ia, n, ia);
          exx += (-0.5 * scl) * d;
          for (int ib = ia + 1; ib < ne; ++ib) {
            cd d2 = 0.0;
            for (int n = 0; n < K; ++n) d2 += Tv(ia, n, ib) * Tv(ib, n, ia);
            exx += (-scl) * d2;
          }
        }
        for (int ia = 0; ia < ne; ++ia)
          for (int n = 0; n < K; ++n) Kl[n] += Tv(ia, n, ia);
      }
      const double sclJ = (walker_type == 1) ? 4.0 : 1.0;
      cd ej = 0.0;
      for (int n = 0; n < K; ++n) ej += Kl[n] * Kl[n];
      E[(long)w * 3 + 0] = e1;
      E[(long)w * 3 + 1] = exx;
      E[(long)w * 3 + 2] = 0.5 * sclJ * ej;
    }
  }
}

// math::exp_hermitian (exp.hpp:27-60): exp(A) for Hermitian A through an eigendecomposition.
// The complex Hermitian A = B + iC is embedded in the real symmetric S = [[B,-C],[C,B]], whose
// matrix exponential is [[Re e^A, -Im e^A],[Im e^A, Re e^A]].
void orc_exp_hermitian(int n, const cd* A, cd* expA) {
  const int m = 2 * n;
  std::vector<double> S((size_t)m * m), w, V;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      cd a = A[(long)i * n + j];
      S[(long)i * m + j] = a.real();
      S[(long)(i + n) * m + j + n] = a.real();
      S[(long)i * m + j + n] = -a.imag();
      S[(long)(i + n) * m + j] = a.imag();
    }
  jacobi_eigh(m, S, w, V);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double re = 0, im = 0;
      for (int k = 0; k < m; ++k) {
        double e = std::exp(w[k]);
        re += V[(long)i * m + k] * e * V[(long)j * m + k];
        im += V[(long)(i + n) * m + k] * e * V[(long)j * m + k];
      }
      expA[(long)i * n + j] = cd(re, im);
    }
}

// P1 as the reference's host path applies it (denseP1 = false, parameters.hpp:120): to_csr
// keeps entries with |x| > 1e-8 (csr_utils.hpp:43-67, AFQMCBasePropagator.cpp:165-166); when
// every kept imaginary part is within 1e-14 of the largest |x|, the real parts are applied in
// real arithmetic, again cut at |x| > 1e-8 (AFQMCBasePropagator.cpp:174-211).
void orc_sparsify_p1(int M, cd* P) {
  const long n = (long)M * M;
  double vmax = 0.0, imax = 0.0;
  for (long e = 0; e < n; ++e) {
    if (std::abs(P[e]) > 1e-8) {
      vmax = std::max(vmax, std::abs(P[e]));
      imax = std::max(imax, std::fabs(P[e].imag()));
    } else {
      P[e] = 0.0;
    }
  }
  if (imax <= 1e-14 * vmax)
    for (long e = 0; e < n; ++e) P[e] = std::fabs(P[e].real()) > 1e-8 ? cd(P[e].real(), 0.0) : cd(0.0);
}

// generateP1 (AFQMCBasePropagator.cpp:85-157) after vMF is known, closed-shell / spin-independent:
//   H1 = dt*(h + v0) + vHS(vMF)   (getOneBodyPropagatorMatrix, Real3IndexFactorization.hpp:99-124,
//                                  add_one_body_shifts one_body.hpp:64-85)
//   H1 <- hermitize(H1) (tensor.hpp:75-93);  P1 = exp_hermitian(-1/2 H1)
void orc_one_body_propagator(int M, int K, double dt, const double* h, const double* v0,
                             const double* Likn, const cd* vMF, cd* P1, cd* H1out) {
  std::vector<cd> ms((size_t)M * M);
  orc_vhs(1, M, K, dt, Likn, vMF, ms.data());
  std::vector<cd> H1((size_t)M * M);
  for (long e = 0; e < (long)M * M; ++e) H1[e] = dt * h[e] + dt * v0[e] + ms[e];
  for (int i = 0; i < M; ++i)
    for (int j = i; j < M; ++j) {
      if (i == j) H1[i * M + i] = H1[i * M + i].real();
      else {
        H1[i * M + j] = 0.5 * H1[i * M + j] + 0.5 * std::conj(H1[j * M + i]);
        H1[j * M + i] = std::conj(H1[i * M + j]);
      }
    }
  if (H1out) std::copy(H1.begin(), H1.end(), H1out);
  for (auto& x : H1) x *= -0.5;
  orc_exp_hermitian(M, H1.data(), P1);
  orc_sparsify_p1(M, P1);
}

// ============================= propagate.hpp ==================================================
// propagate_impl (propagate.hpp:241-298) with apply_expM (41-70), 'N' and a dense P1:
//   Phi <- P1 Phi;  T1 = Phi; for k=1..order: T2 = (i/k) V T1; Phi += T2; swap;  Phi <- P1 Phi
void orc_propagate(int nb, int M, int N, cd* Phi, long ldw, const cd* P1, const cd* V,
                   int order) {
  if (N == 0) return;
#pragma omp parallel
  {
    std::vector<cd> S((size_t)M * N), T1((size_t)M * N), T2((size_t)M * N);
#pragma omp for schedule(static)
    for (int w = 0; w < nb; ++w) {
      cd* P = Phi + (long)w * ldw;
      const cd* Vw = V + (long)w * M * M;
      zgemm_nn(M, N, M, 1.0, P1, M, P, N, 0.0, S.data(), N);
      T1 = S;
      for (int k = 1; k <= order; ++k) {
        cd fact = cd(0.0, 1.0) * (1.0 / double(k));
        zgemm_nn(M, N, M, fact, Vw, M, T1.data(), N, 0.0, T2.data(), N);
        for (size_t e = 0; e < S.size(); ++e) S[e] += T2[e];
        std::swap(T1, T2);
      }
      zgemm_nn(M, N, M, 1.0, P1, M, S.data(), N, 0.0, P, N);
    }
  }
}

// ============================= orthogonalize.hpp ==============================================
// orthogonalize (orthogonalize.hpp:131-169): Householder QR of Phi_w [M,N] with LAPACK zgeqrf /
// zlarfg conventions (beta = -sign(Re alpha) * norm is real), log det R accumulated
// (log_determinant_from_geqrf, determinants_impl.hpp:77-108), Q formed (zungqr) and its columns
// scaled by s_i = sign(Re R_ii) so that the returned Q has a positive diagonal R.
void orc_orthogonalize(int nb, int M, int N, cd* Phi, long ldw, cd* logdetR) {
  if ((long)M * N == 0) return;
  const cd small(std::numeric_limits<double>::min(), 0.0);
#pragma omp parallel
  {
    std::vector<cd> A((size_t)M * N), tau(N), Q((size_t)M * N);
    std::vector<double> scl(N);
#pragma omp for schedule(static)
    for (int w = 0; w < nb; ++w) {
      cd* P = Phi + (long)w * ldw;
      std::copy(P, P + (long)M * N, A.begin());
      const int kmax = std::min(M, N);
      for (int j = 0; j < kmax; ++j) {
        cd alpha = A[(long)j * N + j];
        double xnorm2 = 0.0;
        for (int i = j + 1; i < M; ++i) xnorm2 += std::norm(A[(long)i * N + j]);
        double xnorm = std::sqrt(xnorm2);
        if (xnorm == 0.0 && alpha.imag() == 0.0) {
          tau[j] = 0.0;  // H = I, R_jj = alpha
          continue;
        }
        double nrm = std::sqrt(std::norm(alpha) + xnorm2);
        double beta = (alpha.real() >= 0.0) ? -nrm : nrm;
        tau[j] = cd((beta - alpha.real()) / beta, -alpha.imag() / beta);
        cd sc = 1.0 / (alpha - beta);
        for (int i = j + 1; i < M; ++i) A[(long)i * N + j] *= sc;
        A[(long)j * N + j] = beta;
        // apply H^H = I - conj(tau) v v^H to A[j:, j+1:]
        for (int c = j + 1; c < N; ++c) {
          cd s = A[(long)j * N + c];
          for (int i = j + 1; i < M; ++i) s += std::conj(A[(long)i * N + j]) * A[(long)i * N + c];
          s *= std::conj(tau[j]);
          A[(long)j * N + c] -= s;
          for (int i = j + 1; i < M; ++i) A[(long)i * N + c] -= s * A[(long)i * N + j];
        }
      }
      for (int j = 0; j < kmax; ++j) {
        double r = A[(long)j * N + j].real();
        scl[j] = (r < 0.0) ? -1.0 : 1.0;
        logdetR[w] += std::log(scl[j] * A[(long)j * N + j] + small);
      }
      // Q = H_0 H_1 ... H_{k-1} [I;0], backward accumulation (zung2r)
      std::fill(Q.begin(), Q.end(), cd(0.0));
      for (int i = 0; i < std::min(M, N); ++i) Q[(long)i * N + i] = 1.0;
      for (int j = kmax - 1; j >= 0; --j) {
        for (int c = j; c < N; ++c) {
          cd s = Q[(long)j * N + c];
          for (int i = j + 1; i < M; ++i) s += std::conj(A[(long)i * N + j]) * Q[(long)i * N + c];
          s *= tau[j];
          Q[(long)j * N + c] -= s;
          for (int i = j + 1; i < M; ++i) Q[(long)i * N + c] -= s * A[(long)i * N + j];
        }
      }
      for (int i = 0; i < M; ++i)
        for (int c = 0; c < N; ++c) P[(long)i * N + c] = Q[(long)i * N + c] * (c < kmax ? scl[c] : 1.0);
    }
  }
}

// ============================= Propagator pieces ==============================================
// construct_X_impl (construct_X_kernel.hpp:33-84), continuous charge fields, with the N(0,1)
// field xi given directly in place of probit(U): on entry X holds vbias; on exit
//   X = xi + vdiff, vdiff = i (clamp(vbias) - clamp(vMF)),  HW[w] += -vdiff (X - vdiff/2)
// clamp(z) = z/|z| * bound when |z| > bound; project (cp or force-bias projection) keeps Re.
void orc_construct_X(int W, int K, double vbias_bound, int project_vbias, int free_projection,
                     const cd* vMF, const double* xi, cd* X, cd* HW) {
  const cd im(0.0, 1.0);
#pragma omp parallel for schedule(static)
  for (int w = 0; w < W; ++w) {
    cd hw = 0.0;
    for (int m = 0; m < K; ++m) {
      cd vmf0 = vMF[m];
      cd vmf_t = std::abs(vmf0) > vbias_bound ? vmf0 / std::abs(vmf0) * vbias_bound : vmf0;
      cd x0 = X[(long)w * K + m];
      cd vb_t = std::abs(x0) > vbias_bound ? x0 / std::abs(x0) * vbias_bound : x0;
      if (project_vbias) vb_t = cd(vb_t.real(), 0.0);
      cd vdiff = free_projection ? cd(0.0) : im * (vb_t - vmf_t);
      cd x = xi[(long)w * K + m] + vdiff;
      X[(long)w * K + m] = x;
      hw += -vdiff * (x - 0.5 * vdiff);
    }
    HW[w] = hw;
  }
}

// meanfield_factor (AFQMCBasePropagator.cpp:334-336): mf[w] = -1/2 Re(sum vMF^2) + i sum X vMF
void orc_meanfield_factor(int W, int K, const cd* X, const cd* vMF, cd* mf) {
  cd v2 = 0.0;
  for (int m = 0; m < K; ++m) v2 += vMF[m] * vMF[m];
  for (int w = 0; w < W; ++w) {
    cd s = 0.0;
    for (int m = 0; m < K; ++m) s += X[(long)w * K + m] * vMF[m];
    mf[w] = -0.5 * v2.real() + cd(0.0, 1.0) * s;
  }
}

// Walker property slots after the Slater matrices (WalkerSetBase.cpp:87-117, WalkerConfig.hpp:28-48)
enum { P_WEIGHT = 0, P_PHASE, P_PSEUDO_ELOC, P_E1, P_EXX, P_EJ, P_OVLP, P_LOGSCL_UP, P_LOGSCL_DN,
       P_IS_UNITARY, P_THETA, P_COUNT };

// walker_update (WalkerSetUpdate.hpp:69-191) + bound_local_energy (46-58). stats[0..2] =
// (total, upper, lower) local-energy bound hits; accumulated. With back-propagation the
// reference also records each walker's weight factor (WalkerSetUpdate.hpp:103, 128-147) and its
// updated weight (WalkerSetUpdate.hpp:184-190) in the history rings: wfac_out[W] and
// weight_out[W] receive them when not null.
void orc_walker_update(int W, cd* wbuf, long ldw, long poff, int hybrid, int free_projection,
                       int use_cp, double dt, double Eshift, double E0, const cd* new_ov,
                       const cd* new_E, const cd* mf, const cd* HW, double lower_scale,
                       double upper_scale, long* stats, cd* wfac_out, cd* weight_out) {
  for (int i = 0; i < W; ++i) {
    cd wfac = 0.0;
    cd* p = wbuf + (long)i * ldw + poff;
    cd old_ov = p[P_OVLP], old_eloc = p[P_PSEUDO_ELOC];
    double scale = 1.0, dtheta = 0.0;
    if (!free_projection) {
      dtheta = new_ov[i].imag() - old_ov.imag() - mf[i].imag();
      if (use_cp) scale = std::cos(dtheta) > 0.0 ? 1.0 : 0.0;
      else scale = std::max(0.0, std::cos(dtheta));
    }
    p[P_THETA] = dtheta;
    cd eloc = hybrid ? (mf[i] - HW[i] - (new_ov[i] - old_ov)) / dt + E0
                     : new_E[i * 3 + 0] + new_E[i * 3 + 1] + new_E[i * 3 + 2];
    if (!std::isfinite(eloc.real())) {
      p[P_WEIGHT] = 0.0;
      p[P_PHASE] = 0.0;
      eloc = old_eloc;   // wfac stays 0
    } else {
      if (!free_projection) {
        double hi = Eshift + upper_scale * std::sqrt(2.0 / dt);
        double lo = Eshift - lower_scale * std::sqrt(2.0 / dt);
        stats[0]++;
        if (eloc.real() > hi) stats[1]++;
        else if (eloc.real() < lo) stats[2]++;
        eloc = cd(std::clamp(eloc.real(), lo, hi), eloc.imag());
      }
      cd eeff = hybrid ? eloc : 0.5 * (eloc + old_eloc);
      cd wf = std::exp(-dt * (eeff - Eshift));
      if (std::abs(scale) > std::numeric_limits<double>::min()) wfac = wf / std::abs(wf) / scale;
      p[P_WEIGHT] *= free_projection ? scale * wf : scale * std::abs(wf);
      p[P_PHASE] *= wf / std::abs(wf);
    }
    if (wfac_out) wfac_out[i] = wfac;
    if (weight_out) weight_out[i] = p[P_WEIGHT];
    p[P_PSEUDO_ELOC] = eloc;
    p[P_OVLP] = new_ov[i];
    if (!hybrid) {
      p[P_E1] = new_E[i * 3 + 0];
      p[P_EXX] = new_E[i * 3 + 1];
      p[P_EJ] = new_E[i * 3 + 2];
    }
  }
}

// bound_walker_weights (WalkerSetUpdate.hpp:198-218). stats = (total, upper).
void orc_bound_weights(int W, cd* wbuf, long ldw, long poff, double floor_, double fraction,
                       double target_population, long* stats) {
  const double maxw = std::max(floor_, fraction * target_population);
  for (int i = 0; i < W; ++i) {
    cd* p = wbuf + (long)i * ldw + poff;
    double a = std::abs(p[P_WEIGHT]);
    stats[0]++;
    if (a > maxw) {
      stats[1]++;
      p[P_WEIGHT] *= maxw / a;
    }
  }
}

// comb (population_control.hpp:34-65): offspring counts of W walkers from |w| and one uniform r.
// Returns 0, or 1 when the total weight collapsed (<= 1e-6; the reference aborts there).
int orc_comb(int W, const double* weights, double r, int num_target, int* offspring) {
  std::vector<double> cum(W);
  std::partial_sum(weights, weights + W, cum.begin());
  const double total = cum[W - 1];
  if (!(total > 1e-6)) return 1;
  int last = W - 1;
  while (weights[last] <= 0.0) --last;
  std::fill(offspring, offspring + W, 0);
  int iw = 0, ic = 0;
  while (ic < num_target) {
    const double comb = (ic + r) * total / num_target;
    if (comb < cum[iw] || iw == last) {
      offspring[iw] += 1;
      ic++;
    } else {
      iw++;
    }
  }
  return 0;
}

// population_control (population_control.hpp:219-257) on one rank: comb by |w|, then
// swap_walkers_async's single-rank placement (population_control.hpp:143-209): a walker with k
// copies keeps its slot and its k-1 extra copies fill the slots of zero-offspring walkers in
// increasing slot order; finally every weight becomes w/|w|.
int orc_population_control(int W, cd* wbuf, long ldw, long poff, double r, int* offspring_out) {
  std::vector<double> wt(W);
  for (int i = 0; i < W; ++i) wt[i] = std::abs(wbuf[(long)i * ldw + poff + P_WEIGHT]);
  std::vector<int> off(W);
  if (orc_comb(W, wt.data(), r, W, off.data())) return 1;
  std::vector<int> free_slots;
  for (int i = 0; i < W; ++i)
    if (off[i] == 0) free_slots.push_back(i);
  size_t next = 0;
  for (int i = 0; i < W; ++i)
    for (int c = 1; c < off[i]; ++c) {
      const int dst = free_slots[next++];
      std::copy(wbuf + (long)i * ldw, wbuf + (long)(i + 1) * ldw, wbuf + (long)dst * ldw);
    }
  for (int i = 0; i < W; ++i) {
    cd& w = wbuf[(long)i * ldw + poff + P_WEIGHT];
    w = w / std::abs(w);
  }
  if (offspring_out) std::copy(off.begin(), off.end(), offspring_out);
  return 0;
}

// ============================= whole system (closed-shell NOMSD, one determinant) =============
// A Real3IndexFactorization Hamiltonian + NOMSD trial + AFQMCBasePropagator state: the inputs
// of one walker step, built as RealDenseHamiltonian::getHamiltonianOperations and generateP1 do.
struct orc_system {
  int M, N, K;
  double dt, E0;
  cd ci0;
  std::vector<double> h, Likn, v0;
  std::vector<cd> PsiT, haj, Lnak, vMF, P1, H1;
  // propagator parameters (PropagatorParameters, parameters.hpp:103-126; defaults
  // parameter_defaults.cpp:274-282)
  int taylor_n = 6;
  double vbias_bound = 50.0, upper_cutoff_scale = 10.0, lower_cutoff_scale = 1.0;
  double weight_bound_floor = 100.0, weight_bound_fraction = 0.1;
  int free_projection = 0, use_cp_constraint = 0, project_force_bias = 0;
  long eloc_stats[3] = {0, 0, 0}, weight_stats[2] = {0, 0};
};

orc_system* orc_system_create(int M, int N, int K, const double* h, const double* Likn,
                              const cd* PsiT, double E0, double dt) {
  auto* s = new orc_system();
  s->M = M; s->N = N; s->K = K; s->dt = dt; s->E0 = E0; s->ci0 = 1.0;
  s->h.assign(h, h + (long)M * M);
  s->Likn.assign(Likn, Likn + (long)M * M * K);
  s->PsiT.assign(PsiT, PsiT + (long)N * M);
  s->haj.resize((long)N * M);
  s->Lnak.resize((long)K * N * M);
  s->v0.resize((long)M * M);
  orc_half_rotate(M, N, K, h, Likn, PsiT, s->haj.data(), s->Lnak.data());
  orc_exchange_v0(M, K, Likn, s->v0.data());
  // NOMSD::vMF (NOMSD.cpp:437-465): G of the trial with itself (Phi = PsiT^H), vbias, and
  // math::zero_imag (AFQMCBasePropagator.cpp:107-108).
  std::vector<cd> Phi0((long)M * N), G((long)N * M);
  for (int i = 0; i < M; ++i)
    for (int a = 0; a < N; ++a) Phi0[(long)i * N + a] = std::conj(PsiT[(long)a * M + i]);
  cd ov = 0.0;
  orc_mixed_density_matrix(1, M, N, PsiT, Phi0.data(), 0, G.data(), &ov, 1);
  s->vMF.resize(K);
  orc_vbias(1, M, N, 0, K, 1, dt, s->Lnak.data(), nullptr, G.data(), s->vMF.data());
  for (auto& v : s->vMF) v = cd(v.real(), 0.0);
  s->P1.resize((long)M * M);
  s->H1.resize((long)M * M);
  orc_one_body_propagator(M, K, dt, h, s->v0.data(), Likn, s->vMF.data(), s->P1.data(),
                          s->H1.data());
  return s;
}

void orc_system_destroy(orc_system* s) { delete s; }

void orc_system_set_params(orc_system* s, int taylor_n, double vbias_bound, double upper_scale,
                           double lower_scale, double wfloor, double wfrac, int free_projection,
                           int use_cp, int project_force_bias) {
  s->taylor_n = taylor_n; s->vbias_bound = vbias_bound;
  s->upper_cutoff_scale = upper_scale; s->lower_cutoff_scale = lower_scale;
  s->weight_bound_floor = wfloor; s->weight_bound_fraction = wfrac;
  s->free_projection = free_projection; s->use_cp_constraint = use_cp;
  s->project_force_bias = project_force_bias;
}

// Copies out the derived tensors (any pointer may be null).
void orc_system_tensors(const orc_system* s, cd* haj, cd* Lnak, double* v0, cd* vMF, cd* P1,
                        cd* H1) {
  if (haj) std::copy(s->haj.begin(), s->haj.end(), haj);
  if (Lnak) std::copy(s->Lnak.begin(), s->Lnak.end(), Lnak);
  if (v0) std::copy(s->v0.begin(), s->v0.end(), v0);
  if (vMF) std::copy(s->vMF.begin(), s->vMF.end(), vMF);
  if (P1) std::copy(s->P1.begin(), s->P1.end(), P1);
  if (H1) std::copy(s->H1.begin(), s->H1.end(), H1);
}

// NOMSD::MixedDensityMatrix / DensityMatrix, CLOSED, ndet=1 (NOMSD.cpp:219-279, NOMSD.hpp:292-331):
// compact G and ov = 2 * logdet.
void orc_system_mixed_density_matrix(const orc_system* s, int W, const cd* wbuf, long ldw, cd* G,
                                     cd* ov) {
  for (int w = 0; w < W; ++w) ov[w] = 0.0;
  orc_mixed_density_matrix(W, s->M, s->N, s->PsiT.data(), wbuf, ldw, G, ov, 1);
  for (int w = 0; w < W; ++w) ov[w] *= 2.0;
}

// NOMSD::Log_Overlap (NOMSD.cpp:280-301), CLOSED ndet=1: ov = 2 logdet + log(conj c0).
void orc_system_log_overlap(const orc_system* s, int W, const cd* wbuf, long ldw, cd* ov) {
  for (int w = 0; w < W; ++w) ov[w] = 0.0;
  orc_log_overlap(W, s->M, s->N, s->PsiT.data(), wbuf, ldw, ov);
  cd val = std::log(std::abs(s->ci0) < 1e-12 ? cd(1e-12) : std::conj(s->ci0));
  for (int w = 0; w < W; ++w) ov[w] = 2.0 * ov[w] + val;
}

// NOMSD::Energy (NOMSD.cpp:159-182), CLOSED ndet=1: E[w] = (E1,EXX,EJ), ov = 2 logdet + log conj c0.
void orc_system_energy(const orc_system* s, int W, const cd* wbuf, long ldw, cd* E, cd* ov) {
  std::vector<cd> G((long)W * s->N * s->M);
  orc_system_mixed_density_matrix(s, W, wbuf, ldw, G.data(), ov);
  orc_energy(W, s->M, s->N, 0, s->K, 1, s->E0, s->haj.data(), nullptr, s->Lnak.data(), nullptr,
             G.data(), E);
  cd val = std::log(std::abs(s->ci0) < 1e-12 ? cd(1e-12) : std::conj(s->ci0));
  for (int w = 0; w < W; ++w) ov[w] += val;
}

// Seeds the walker properties as execute_simulation does before the first step
// (execute.cpp:172-184 via NOMSD::Energy(wset) NOMSD.cpp:69-80 and seedPseudoEnergy).
void orc_system_seed_properties(const orc_system* s, int W, cd* wbuf, long ldw) {
  const long poff = (long)s->M * s->N;
  std::vector<cd> E((long)W * 3), ov(W);
  orc_system_energy(s, W, wbuf, ldw, E.data(), ov.data());
  for (int w = 0; w < W; ++w) {
    cd* p = wbuf + (long)w * ldw + poff;
    p[P_OVLP] = ov[w];
    p[P_E1] = E[w * 3 + 0];
    p[P_EXX] = E[w * 3 + 1];
    p[P_EJ] = E[w * 3 + 2];
    p[P_PSEUDO_ELOC] = E[w * 3 + 0] + E[w * 3 + 1] + E[w * 3 + 2];
  }
}

// AFQMCBasePropagator::Propagate (AFQMCBasePropagator.cpp:289-392) for one step, CLOSED,
// continuous charge fields, dense vHS, with the N(0,1) fields xi[W][K] given. Optional outputs
// (may be null) expose the intermediates for step-level parity: vbias, X, V, new overlaps/energies.
void orc_system_step(orc_system* s, int W, cd* wbuf, long ldw, const double* xi, double Eshift,
                     int hybrid, cd* out_vbias, cd* out_X, cd* out_V, cd* out_ov, cd* out_E,
                     cd* out_wfac, cd* out_weight) {
  const int M = s->M, N = s->N, K = s->K;
  const double dt = s->dt;
  const long poff = (long)M * N;
  std::vector<cd> G((long)W * N * M), ov(W), X((long)W * K), HW(W), mf(W), V((long)W * M * M),
      new_ov(W, 0.0), new_E((long)W * 3, 0.0);
  // 1. vbias (NOMSD::vbias, NOMSD.cpp:39-61)
  orc_system_mixed_density_matrix(s, W, wbuf, ldw, G.data(), ov.data());
  orc_vbias(W, M, N, 0, K, 1, dt, s->Lnak.data(), nullptr, G.data(), X.data());
  if (out_vbias) std::copy(X.begin(), X.end(), out_vbias);
  // 2. assemble_X (AFQMCBasePropagator.cpp:686-733) + meanfield factor (334-336)
  orc_construct_X(W, K, s->vbias_bound, s->use_cp_constraint || s->project_force_bias,
                  s->free_projection, s->vMF.data(), xi, X.data(), HW.data());
  orc_meanfield_factor(W, K, X.data(), s->vMF.data(), mf.data());
  if (out_X) std::copy(X.begin(), X.end(), out_X);
  // 3. vHS, 4. propagate
  orc_vhs(W, M, K, dt, s->Likn.data(), X.data(), V.data());
  if (out_V) std::copy(V.begin(), V.end(), out_V);
  orc_propagate(W, M, N, wbuf, ldw, s->P1.data(), V.data(), s->taylor_n);
  // 5. overlap (hybrid) or energy
  if (hybrid) orc_system_log_overlap(s, W, wbuf, ldw, new_ov.data());
  else orc_system_energy(s, W, wbuf, ldw, new_E.data(), new_ov.data());
  if (out_ov) std::copy(new_ov.begin(), new_ov.end(), out_ov);
  if (out_E) std::copy(new_E.begin(), new_E.end(), out_E);
  // 6. walker_update, 7. bound_walker_weights (target population = W on one rank)
  orc_walker_update(W, wbuf, ldw, poff, hybrid, s->free_projection, s->use_cp_constraint, dt,
                    Eshift, s->E0, new_ov.data(), new_E.data(), mf.data(), HW.data(),
                    s->lower_cutoff_scale, s->upper_cutoff_scale, s->eloc_stats, out_wfac,
                    out_weight);
  orc_bound_weights(W, wbuf, ldw, poff, s->weight_bound_floor, s->weight_bound_fraction,
                    double(W), s->weight_stats);
}

// AFQMCBasePropagator::Orthogonalize (AFQMCBasePropagator.cpp:735-741): QR, then OVLP recomputed.
void orc_system_orthogonalize(const orc_system* s, int W, cd* wbuf, long ldw) {
  std::vector<cd> ldet(W, 0.0), ov(W);
  orc_orthogonalize(W, s->M, s->N, wbuf, ldw, ldet.data());
  orc_system_log_overlap(s, W, wbuf, ldw, ov.data());
  const long poff = (long)s->M * s->N;
  for (int w = 0; w < W; ++w) wbuf[(long)w * ldw + poff + P_OVLP] = ov[w];
}

void orc_system_stats(const orc_system* s, long* eloc_stats3, long* weight_stats2) {
  for (int i = 0; i < 3; ++i) eloc_stats3[i] = s->eloc_stats[i];
  for (int i = 0; i < 2; ++i) weight_stats2[i] = s->weight_stats[i];
}

// ---------------------------------------------------------------------------------------------
// Particle-hole multi-determinant (PHMSD) energy corrections of the excited determinants.
//
// ph_excited_2body_energy_dense_cholesky (detail/phmsd_impl.hpp:29-147, HOST branch): for each
// walker w and excited determinant d, whose excitation iexcit[d] = (electrons p_1..p_nex, i.e.
// rows of refc, then the orbitals v_1..v_nex they move to), with the caller's walker tensor
// Twina[w][i][n][a] (electron i, Cholesky vector n, active orbital a) and the compact R[w][d][e][a]
// of the determinant:
//   occ(i)  = refc(i), or v_e where i = p_e
//   F_e     = Twina[w][p_e] R[w][d][e]                      (gemv over a, a vector over n)
//   eX      = sum_e sum_n F_e^2 + 2 sum_{e<f} sum_n (Twina[p_f] R_e)(Twina[p_e] R_f)
//           + 2 sum_e sum_j sum_a (sum_n Twina[j][n][a] Twina[p_e][n][occ j]) R_e[a]
//           + sum_i (sum_n T[i][n][occ i]^2 - T[i][n][ref i]^2)         over changed rows
//           + 2 sum_{i<j} (sum_n T[i][n][occ j] T[j][n][occ i] - T[i][n][ref j] T[j][n][ref i])
//             where row i or j changed
//   KE[d][w][n] = sum_e F_e[n] + sum_i Twina[w][i][n][occ i]
//   EX[w] -= wgt[d][w] eX / 2;  EJ[w] += wgt[d][w] (sum_n KE^2 - (sum_i T[i][n][ref i])^2) / 2
// All products are non-conjugated (complex-symmetric).
void orc_ph_excited_2body(int nw, int ndet, int nex, int nact, int nelec, int nchol,
                          const int* iexcit, const int* refc, const cd* Twina, const cd* R,
                          const cd* wgt, cd* EX, cd* EJ, cd* KE) {
#pragma omp parallel for schedule(dynamic)
  for (int iw = 0; iw < nw; ++iw) {
    std::vector<int> occ(nelec);
    std::vector<cd> Fn(nchol), Fn1(nchol), Fa(nact), Kn(nchol), F0(nchol);
    auto T = [&](int i, int n, int a) -> cd {
      return Twina[(((long)iw * nelec + i) * nchol + n) * nact + a];
    };
    // T[i] (nchol x nact) times a vector over a
    auto gemv = [&](int i, const cd* x, cd* y) {
      for (int n = 0; n < nchol; ++n) {
        cd acc = 0.0;
        for (int a = 0; a < nact; ++a) acc += T(i, n, a) * x[a];
        y[n] = acc;
      }
    };
    cd ex_w = 0.0, ej_w = 0.0;
    // reference part of the Coulomb term: F0 = sum_i T[i][:][refc i]
    for (int n = 0; n < nchol; ++n) {
      cd acc = 0.0;
      for (int i = 0; i < nelec; ++i) acc += T(i, n, refc[i]);
      F0[n] = acc;
    }
    cd eJ0 = 0.0;
    for (int n = 0; n < nchol; ++n) eJ0 += F0[n] * F0[n];
    for (int d = 0; d < ndet; ++d) {
      const int* ip = iexcit + (long)d * 2 * nex;     // occupied (particles removed)
      const int* iv = ip + nex;                        // virtual (added)
      const cd* Rd = R + (((long)iw * ndet + d) * nex) * nact;
      for (int i = 0; i < nelec; ++i) {
        occ[i] = refc[i];
        for (int e = 0; e < nex; ++e)
          if (ip[e] == i) { occ[i] = iv[e]; break; }
      }
      std::fill(Kn.begin(), Kn.end(), cd(0.0));
      cd eX = 0.0;
      for (int e = 0; e < nex; ++e) {
        gemv(ip[e], Rd + (long)e * nact, Fn.data());
        for (int n = 0; n < nchol; ++n) {
          eX += Fn[n] * Fn[n];
          Kn[n] += Fn[n];
        }
        for (int f = e + 1; f < nex; ++f) {
          gemv(ip[f], Rd + (long)e * nact, Fn.data());
          gemv(ip[e], Rd + (long)f * nact, Fn1.data());
          cd acc = 0.0;
          for (int n = 0; n < nchol; ++n) acc += Fn[n] * Fn1[n];
          eX += 2.0 * acc;
        }
        for (int j = 0; j < nelec; ++j) {
          const int oj = occ[j];
          for (int a = 0; a < nact; ++a) {
            cd acc = 0.0;
            for (int n = 0; n < nchol; ++n) acc += T(j, n, a) * T(ip[e], n, oj);
            Fa[a] = acc;
          }
          cd acc = 0.0;
          for (int a = 0; a < nact; ++a) acc += Fa[a] * Rd[(long)e * nact + a];
          eX += 2.0 * acc;
        }
      }
      for (int i = 0; i < nelec; ++i) {
        const int oi = occ[i], ri = refc[i];
        if (oi != ri) {
          cd acc = 0.0;
          for (int n = 0; n < nchol; ++n) acc += T(i, n, oi) * T(i, n, oi) - T(i, n, ri) * T(i, n, ri);
          eX += acc;
        }
        for (int j = i + 1; j < nelec; ++j) {
          const int oj = occ[j], rj = refc[j];
          if (oi != ri || oj != rj) {
            cd acc = 0.0;
            for (int n = 0; n < nchol; ++n)
              acc += T(i, n, oj) * T(j, n, oi) - T(i, n, rj) * T(j, n, ri);
            eX += 2.0 * acc;
          }
        }
      }
      for (int i = 0; i < nelec; ++i)
        for (int n = 0; n < nchol; ++n) Kn[n] += T(i, n, occ[i]);
      cd kk = 0.0;
      for (int n = 0; n < nchol; ++n) {
        kk += Kn[n] * Kn[n];
        KE[((long)d * nw + iw) * nchol + n] = Kn[n];
      }
      const cd wd = wgt[(long)d * nw + iw];
      ex_w -= 0.5 * wd * eX;
      ej_w += 0.5 * wd * (kk - eJ0);
    }
    EX[iw] += ex_w;
    EJ[iw] += ej_w;
  }
}

// ph_excited_1body_energy (detail/phmsd_impl.hpp:149-183, HOST branch):
//   E[w] += sum_d wgt[d][w] sum_e ( sum_a S[w][p_e][a] R[w][d][e][a] + S[w][p_e][v_e] - S[w][p_e][refc p_e] )
void orc_ph_excited_1body(int nw, int ndet, int nex, int nact, int nelec, const int* iexcit,
                          const int* refc, const cd* S, const cd* R, const cd* wgt, cd* E) {
  for (int iw = 0; iw < nw; ++iw) {
    cd acc_w = 0.0;
    for (int d = 0; d < ndet; ++d) {
      const int* ip = iexcit + (long)d * 2 * nex;
      const int* iv = ip + nex;
      cd e_ = 0.0;
      for (int e = 0; e < nex; ++e) {
        const cd* Sp = S + ((long)iw * nelec + ip[e]) * nact;
        const cd* Re = R + (((long)iw * ndet + d) * nex + e) * nact;
        for (int a = 0; a < nact; ++a) e_ += Sp[a] * Re[a];
        e_ += Sp[iv[e]] - Sp[refc[ip[e]]];
      }
      acc_w += wgt[(long)d * nw + iw] * e_;
    }
    E[iw] += acc_w;
  }
}

// ---------------------------------------------------------------- PHMSD overlaps and R
// One rank-nex shell: e = iexcit[d] (electrons e_0..e_{nex-1}, target orbitals after them);
// T [nw][nact][nelec]; A(p, q) = T[w][e_{nex+p}][e_q].
static bool ph_factor(int nex, int nact, int nelec, const int* e, const cd* Tw, cd* Q,
                      int* ipiv, cd& ov) {
  for (int p = 0; p < nex; ++p)
    for (int q = 0; q < nex; ++q) Q[p * nex + q] = Tw[(long)e[nex + p] * nelec + e[q]];
  getrf(nex, Q, ipiv);
  log_det_from_getrf(nex, Q, ipiv, ov);     // ov = log det (reference: exp(log det))
  ov = std::exp(ov);
  return std::abs(ov) != 0.0;
}

// MixedDensityMatrixForWoodbury (density_matrix.hpp:651-717), compact: TAB = A B_w,
// TNN(i, :) = TAB(ref(i), :), getrf, ov[w] += log det (log_determinant_from_getrf),
// getri_or_zero, QQ0 = TAB inv(TNN), G = inv(TNN)^T B_w^T
void orc_ph_woodbury(int nb, int nmo, int nel, int nact, const cd* A, const int* ref,
                     const cd* Phi, long ldw, cd* G, cd* ov, cd* QQ0) {
  for (int w = 0; w < nb; ++w) {
    const cd* B = Phi + (long)w * ldw;                    // [nmo][nel]
    std::vector<cd> TAB((long)nact * nel, 0.0), TNN((long)nel * nel);
    for (int a = 0; a < nact; ++a)
      for (int k = 0; k < nmo; ++k) {
        const cd x = A[(long)a * nmo + k];
        for (int e = 0; e < nel; ++e) TAB[(long)a * nel + e] += x * B[(long)k * nel + e];
      }
    for (int i = 0; i < nel; ++i)
      for (int e = 0; e < nel; ++e) TNN[(long)i * nel + e] = TAB[(long)ref[i] * nel + e];
    std::vector<int> ipiv(nel);
    getrf(nel, TNN.data(), ipiv.data());
    log_det_from_getrf(nel, TNN.data(), ipiv.data(), ov[w]);
    getri_or_zero(nel, TNN.data(), ipiv.data());
    for (int a = 0; a < nact; ++a)
      for (int e = 0; e < nel; ++e) {
        cd acc = 0.0;
        for (int j = 0; j < nel; ++j) acc += TAB[(long)a * nel + j] * TNN[(long)j * nel + e];
        QQ0[((long)w * nact + a) * nel + e] = acc;
      }
    for (int e = 0; e < nel; ++e)
      for (int k = 0; k < nmo; ++k) {
        cd acc = 0.0;
        for (int j = 0; j < nel; ++j) acc += TNN[(long)j * nel + e] * B[(long)k * nel + j];
        G[((long)w * nel + e) * nmo + k] = acc;
      }
  }
}

// PHMSD.cpp:441-450: W = csrmm(C, ratio_dn); Ov = contract(ratio_up, W); wgt = W * ratio_up
void orc_ph_spin_weights(int nw, int nu, const int* cp, const int* ci, const cd* cv,
                         const cd* rup, const cd* rdn, cd* wgt, cd* ov) {
  for (int u = 0; u < nu; ++u)
    for (int w = 0; w < nw; ++w) {
      cd acc = 0.0;
      for (int j = cp[u]; j < cp[u + 1]; ++j) acc += cv[j] * rdn[(long)ci[j] * nw + w];
      wgt[(long)u * nw + w] = acc;
    }
  for (int w = 0; w < nw; ++w) {
    cd acc = 0.0;
    for (int u = 0; u < nu; ++u) acc += rup[(long)u * nw + w] * wgt[(long)u * nw + w];
    ov[w] = acc;
  }
  for (long i = 0; i < (long)nu * nw; ++i) wgt[i] *= rup[i];
}

// calculate_overlaps (phmsd_helpers.hpp:111-195), one shell: ov[d][w] = det A
void orc_ph_overlaps(int nw, int ndet, int nex, int nact, int nelec, const int* iexcit,
                     const cd* T, cd* ov) {
  for (int d = 0; d < ndet; ++d)
    for (int w = 0; w < nw; ++w) {
      std::vector<cd> Q(nex * nex);
      std::vector<int> ipiv(nex);
      cd o;
      ph_factor(nex, nact, nelec, iexcit + (long)d * 2 * nex, T + (long)w * nact * nelec,
                Q.data(), ipiv.data(), o);
      // a zero pivot gives log(DBL_MIN)-sized values in log_det; the determinant itself is 0
      bool sing = false;
      for (int j = 0; j < nex; ++j) sing |= Q[j * nex + j] == cd(0.0);
      ov[(long)d * nw + w] = sing ? cd(0.0) : o;
    }
}

// get_compact_ph_R_matrices (phmsd_helpers.hpp:317-383): R [nw][ndet][nex][nact], zeroed first
void orc_ph_compact_R(int nw, int ndet, int nex, int nact, int nelec, const int* iexcit,
                      const int* refc, const cd* T, cd* R) {
  std::fill(R, R + (long)nw * ndet * nex * nact, cd(0.0));
  std::vector<cd> Q(nex * nex);
  std::vector<int> ipiv(nex), orbs(nelec);
  for (int w = 0; w < nw; ++w) {
    const cd* Tw = T + (long)w * nact * nelec;
    for (int d = 0; d < ndet; ++d) {
      const int* e = iexcit + (long)d * 2 * nex;
      cd* Rd = R + ((long)w * ndet + d) * nex * nact;
      for (int i = 0; i < nelec; ++i) {
        orbs[i] = refc[i];
        for (int p = 0; p < nex; ++p)
          if (e[p] == i) {
            orbs[i] = e[nex + p];
            break;
          }
      }
      cd ov;
      ph_factor(nex, nact, nelec, e, Tw, Q.data(), ipiv.data(), ov);
      bool sing = false;
      for (int j = 0; j < nex; ++j) sing |= Q[j * nex + j] == cd(0.0);
      if (sing) continue;
      getri_or_zero(nex, Q.data(), ipiv.data());
      for (int p = 0; p < nex; ++p)
        for (int q = 0; q < nex; ++q) {
          const cd Ipq = Q[p * nex + q];
          const cd* Tq = Tw + (long)e[q + nex] * nelec;
          for (int i = 0; i < nelec; ++i) Rd[(long)p * nact + orbs[i]] -= Ipq * Tq[i];
          Rd[(long)p * nact + orbs[e[q]]] += Ipq;
        }
    }
  }
}

// calculate_R (phmsd_helpers.hpp:200-262), host path, for shells nex = 1..nshell with the
// excitation lists concatenated (shell k holds ndets[k] determinants, 2 k ints each) and the
// weights wgt [1 + sum ndets][nw] (row 0 the reference): R [nw][nelec][nact], zeroed first
void orc_ph_calculate_R(int nw, int nshell, const int* ndets, int nact, int nelec,
                        const int* iexcit, const int* refc, const cd* T, const cd* wgt, cd* R) {
  std::fill(R, R + (long)nw * nelec * nact, cd(0.0));
  for (int i = 0; i < nelec; ++i)
    for (int w = 0; w < nw; ++w) R[((long)w * nelec + i) * nact + refc[i]] += wgt[w];
  std::vector<int> orbs(nelec);
  const int* ex = iexcit;
  for (int nex = 1, nd = 1; nex <= nshell; ++nex) {
    std::vector<cd> Q(nex * nex);
    std::vector<int> ipiv(nex);
    for (int k = 0; k < ndets[nex - 1]; ++k, ++nd, ex += 2 * nex) {
      for (int i = 0; i < nelec; ++i) {
        orbs[i] = refc[i];
        for (int p = 0; p < nex; ++p)
          if (ex[p] == i) {
            orbs[i] = ex[nex + p];
            break;
          }
      }
      for (int w = 0; w < nw; ++w) {
        const cd* Tw = T + (long)w * nact * nelec;
        cd ov;
        ph_factor(nex, nact, nelec, ex, Tw, Q.data(), ipiv.data(), ov);
        bool sing = false;
        for (int j = 0; j < nex; ++j) sing |= Q[j * nex + j] == cd(0.0);
        if (sing) continue;
        getri_or_zero(nex, Q.data(), ipiv.data());
        const cd wd = wgt[(long)nd * nw + w];
        cd* Rw = R + (long)w * nelec * nact;
        for (int i = 0; i < nelec; ++i) Rw[(long)i * nact + orbs[i]] += wd;
        for (int p = 0; p < nex; ++p) {
          cd* Rp = Rw + (long)ex[p] * nact;
          for (int q = 0; q < nex; ++q) {
            const cd Ipq = Q[p * nex + q];
            const cd* Tq = Tw + (long)ex[q + nex] * nelec;
            for (int i = 0; i < nelec; ++i) Rp[orbs[i]] -= wd * Ipq * Tq[i];
            Rp[orbs[ex[q]]] += wd * Ipq;
          }
        }
      }
    }
  }
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
