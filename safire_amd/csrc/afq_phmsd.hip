// The particle-hole multi-determinant (PHMSD) trial over COLLINEAR walkers: its construction from
// a CI expansion and its local energy, on the C ABI.
//
//   afq_ph_create  = phmsd_from_params (Wavefunction.cpp:341-497) with build_ph_struct
//                    (readWfn.cpp:218-340): the reference is determinant 0; every other
//                    determinant is written per spin as an excitation of it (positions of the
//                    excited reference orbitals, then the orbitals they go to, in the order the
//                    determinant lists them), unique per spin and rank in order of first
//                    appearance, with the CI coefficient's sign fixed for the in-place order;
//                    every orbital of the basis is active (the counts of find_active_space,
//                    Excitations.hpp:71-102, are never negative), so orbital i and i + nmo map
//                    to active index i; the couplings conj(ci) go into CSR matrices [nu][nd] and
//                    [nd][nu] (Wavefunction.cpp:440-466). The half-rotated tensors of each spin
//                    (haj = A h, Lna_n = A L_n, A the trial orbitals as rows) are built on the GPU.
//   afq_ph_energy  = PHMSD::energy_alg1 (PHMSD.cpp:377-544), in walker blocks (the reference's
//                    nwalk_block_size) and determinant blocks (ndet_block_size), both
//                    numerically irrelevant: every kernel is the GPU form of one reference call
//                    (k_phmsd.hip), in the reference's order.
//   afq_ph_mixed_density_matrix = PHMSD::MixedDensityMatrix (PHMSD.cpp:549-795), compact or full.
//   afq_ph_vmf     = PHMSD::vMF (PHMSD.cpp:1329-1565): the trial's mean-field density on the host
//                    (as the reference's host path forms it), then vbias on the GPU.
//   afq_ph_vbias   = PHMSD::vbias (PHMSD.cpp:53-72): the compact G of both spins contracted with
//                    the half-rotated Cholesky vectors of both spins stacked per vector
//                    (Real3IndexFactorization.hpp:328-361, sqrt(dt) for COLLINEAR walkers).
#include "afq_internal.hpp"

#include <algorithm>
#include <complex>
#include <memory>
#include <new>
#include <vector>

namespace afq {

void launch_half_rotate(hipStream_t s, int M, int N, int K, const double* PsiT, const double* h,
                        const double* Likn, double* haj, double* Lnak);

namespace {

constexpr int kPhDetBlock = 256;   // determinants per R block

struct Shell {
  int nex = 0, ndet = 0, row = 0;  // rank, unique excitations, first row in the ratio arrays
  int* iexcit = nullptr;           // device [ndet][2 nex]
};

struct Spin {
  int nel = 0, nunique = 1;
  std::vector<std::vector<int>> confs;          // host: occupation of every unique excitation
  int* dconf = nullptr;            // [nunique][nel]: confs on the GPU
  std::vector<int> hcp, hci;                    // host copy of the couplings (CSR)
  std::vector<std::complex<double>> hcv;
  std::vector<std::complex<double>> hA;         // host copy of the orbital rows [nact][nmo]
  double* A = nullptr;             // c128 [nact][nmo]
  int* refc = nullptr;             // [nel]
  double* Lna = nullptr;           // c128 [nchol][nact][nmo]
  double* haj = nullptr;           // c128 [nact][nmo]
  std::vector<Shell> shells;
  int *cp = nullptr, *ci = nullptr;  // couplings of this spin's excitations with the other's
  double* cv = nullptr;
};

template <typename T>
T* dev_alloc(std::vector<void*>& owned, size_t n) {
  void* p = nullptr;
  AFQ_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  owned.push_back(p);
  return static_cast<T*>(p);
}

template <typename T>
T* dev_copy(std::vector<void*>& owned, const T* src, size_t n) {
  T* p = dev_alloc<T>(owned, n);
  if (n) AFQ_HIP(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

// get_excitation_number (Excitations.hpp:148-192) for sorted ref and conf: positions of the
// reference orbitals missing from conf, the orbitals of conf missing from ref, and the sign of
// sorting ref with the first replaced in place by the second (exchange count of a bubble pass).
int excitation_of(const std::vector<int>& ref, const std::vector<int>& conf, std::vector<int>& ex) {
  ex.clear();
  std::vector<int> parts, work(ref);
  for (size_t i = 0; i < ref.size(); ++i)
    if (!std::binary_search(conf.begin(), conf.end(), ref[i])) ex.push_back((int)i);
  for (int o : conf)
    if (!std::binary_search(ref.begin(), ref.end(), o)) parts.push_back(o);
  AFQ_CHECK(parts.size() == ex.size(), AFQ_EINVAL,
            "afq_ph_create: a determinant has a different electron count than the reference");
  for (size_t c = 0; c < parts.size(); ++c) work[ex[c]] = parts[c];
  int sign = 1;
  for (size_t i = 0; i < work.size(); ++i)
    for (size_t j = i + 1; j < work.size(); ++j)
      if (work[j] < work[i]) {
        sign = -sign;
        std::swap(work[i], work[j]);
      }
  ex.insert(ex.end(), parts.begin(), parts.end());
  return sign;
}

using zc = std::complex<double>;

// LU with partial pivoting of a [n][n] row-major matrix in place; returns log det (the sign of
// the row exchanges included) and leaves the inverse in inv when it is given
zc host_logdet_inverse(std::vector<zc> a, int n, std::vector<zc>* inv) {
  std::vector<int> piv(n);
  zc ld = 0.0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    for (int r = j + 1; r < n; ++r)
      if (std::abs(a[r * n + j]) > std::abs(a[p * n + j])) p = r;
    piv[j] = p;
    if (p != j) {
      for (int c = 0; c < n; ++c) std::swap(a[j * n + c], a[p * n + c]);
      ld += zc(0.0, M_PI);
    }
    const zc d = a[j * n + j];
    ld += std::log(d);
    for (int r = j + 1; r < n; ++r) {
      const zc f = a[r * n + j] / d;
      a[r * n + j] = f;
      for (int c = j + 1; c < n; ++c) a[r * n + c] -= f * a[j * n + c];
    }
  }
  if (inv) {
    inv->assign((size_t)n * n, 0.0);
    for (int col = 0; col < n; ++col) {
      std::vector<zc> b(n, 0.0);
      b[col] = 1.0;
      for (int j = 0; j < n; ++j) std::swap(b[j], b[piv[j]]);
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < r; ++c) b[r] -= a[r * n + c] * b[c];
      for (int r = n - 1; r >= 0; --r) {
        for (int c = r + 1; c < n; ++c) b[r] -= a[r * n + c] * b[c];
        b[r] /= a[r * n + r];
      }
      for (int r = 0; r < n; ++r) (*inv)[(size_t)r * n + col] = b[r];
    }
  }
  return ld;
}

}  // namespace

}  // namespace afq

struct afq_ph_t {
  int nmo = 0, nchol = 0, nact = 0, wblock = 0, device = 0;
  int energy_alg = AFQ_PH_ENERGY_WOODBURY;
  double E0 = 0.0;
  afq::Spin spin[2];
  double* Lstack = nullptr;        // c128 [nchol][2 nact][nmo]: both spins' Lna per vector
  double* h = nullptr;             // f64 [nmo][nmo]
  double* Likn = nullptr;          // f64 [nmo * nmo][nchol]
  // propagator state (afq_ph_prepare) and the step's per-walker workspace
  double dt = 0.0;
  afq_step_params params;
  double *vMF = nullptr, *v0 = nullptr, *H1 = nullptr, *P1 = nullptr;
  int wcap = 0;
  double *X = nullptr, *HW = nullptr, *mf = nullptr, *V = nullptr, *ovw = nullptr, *Ew = nullptr,
         *cum = nullptr;
  int *off = nullptr, *src = nullptr, *ok = nullptr;
  unsigned long long* stats = nullptr;
  std::vector<void*> step_owned;
  std::vector<void*> owned;
  // workspace for one walker block
  double *G[2] = {nullptr, nullptr}, *QQ0[2] = {nullptr, nullptr}, *work = nullptr;
  double *ratios[2] = {nullptr, nullptr}, *wgt[2] = {nullptr, nullptr};
  double *Ov = nullptr, *Ov2 = nullptr, *log_ov = nullptr, *Eref = nullptr, *Ex = nullptr,
         *acc = nullptr;
  double *Twina = nullptr, *Swia = nullptr, *KEa = nullptr, *KEb = nullptr, *KEl = nullptr,
         *R = nullptr;
  double *Racc = nullptr, *Gblk = nullptr, *split = nullptr, *p2ws = nullptr;
  double* Emsd = nullptr;          // c128 [2][max nunique][wblock][4] (energy_alg0)
  double* Tinv0 = nullptr;         // c128 [max nunique][wblock][nel][nel] (energy_alg0)
  long split_elems = 0, p2ws_doubles = 0;
  ~afq_ph_t() {
    for (void* p : owned) (void)hipFree(p);
    for (void* p : step_owned) (void)hipFree(p);
  }
};

using namespace afq;

namespace {

template <typename F>
int ph_guard(F&& f) {
  try {
    f();
    return AFQ_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("out of host memory");
    return AFQ_ENOMEM;
  } catch (...) {
    set_error("unknown error");
    return AFQ_EINVAL;
  }
}

// Step 1 of energy_alg1 / MixedDensityMatrix: the reference G, QQ0 and log overlap of both spins
// and the excitation overlap ratios (row 0 = the reference).
void ph_overlaps_block(afq_ph_t* ph, hipStream_t s, int nb, const double* const Phi[2], long ldw) {
  Spin* S = ph->spin;
  launch_ph_fill(s, nb, 0.0, 0.0, ph->log_ov);
  for (int sp = 0; sp < 2; ++sp) {
    launch_ph_woodbury(s, nb, ph->nmo, S[sp].nel, ph->nact, S[sp].A, S[sp].refc, Phi[sp], ldw,
                       ph->G[sp], ph->log_ov, ph->QQ0[sp], ph->work);
    launch_ph_fill(s, nb, 1.0, 0.0, ph->ratios[sp]);
    for (const Shell& sh : S[sp].shells)
      launch_ph_overlaps(s, nb, sh.ndet, sh.nex, ph->nact, S[sp].nel, sh.iexcit, ph->QQ0[sp],
                         ph->ratios[sp] + 2L * sh.row * nb);
  }
}

// PHMSD::MixedDensityMatrix (PHMSD.cpp:692-795), compact, for one walker block: G rows
// [0, nact) and [nact, 2 nact) of d_Gc [nb][2 nact][nmo]; d_logov [nb] = log Ov + log_ov.
void ph_density_block(afq_ph_t* ph, hipStream_t s, int nb, const double* const Phi[2], long ldw,
                      double* d_Gc, double* d_logov) {
  Spin* S = ph->spin;
  const int na = ph->nact;
  ph_overlaps_block(ph, s, nb, Phi, ldw);
  for (int sp = 0; sp < 2; ++sp) {
    // weights W[u] = ratio_s[u] sum_d C[u][d] ratio_other[d] (the up spin's also give Ov)
    launch_ph_spin_weights(s, nb, S[sp].nunique, S[sp].cp, S[sp].ci, S[sp].cv, ph->ratios[sp],
                           ph->ratios[1 - sp], ph->wgt[sp], sp == 0 ? ph->Ov : ph->Ov2);
    // calculate_R: the reference, then every shell in determinant blocks
    launch_ph_fill(s, (long)nb * S[sp].nel * na, 0.0, 0.0, ph->Racc);
    launch_ph_accumulate_R(s, nb, 1, 0, na, S[sp].nel, nullptr, S[sp].refc, nullptr, nullptr,
                           ph->wgt[sp], ph->Racc);
    for (const Shell& sh : S[sp].shells)
      for (int d0 = 0; d0 < sh.ndet; d0 += kPhDetBlock) {
        const int nbk = std::min(kPhDetBlock, sh.ndet - d0);
        const int* iex = sh.iexcit + 2L * sh.nex * d0;
        launch_ph_compact_R(s, nb, nbk, sh.nex, na, S[sp].nel, iex, S[sp].refc, ph->QQ0[sp],
                            ph->R);
        launch_ph_accumulate_R(s, nb, nbk, sh.nex, na, S[sp].nel, iex, S[sp].refc, ph->R,
                               ph->ratios[sp] + 2L * (sh.row + d0) * nb,
                               ph->wgt[sp] + 2L * (sh.row + d0) * nb, ph->Racc);
      }
    launch_ph_rg(s, nb, S[sp].nel, na, ph->nmo, 2 * na, sp * na, ph->Racc, ph->G[sp], ph->Ov,
                 d_Gc);
  }
  launch_ph_log_overlap(s, nb, ph->Ov, ph->log_ov, d_logov);
}

// PHMSD::energy_alg1 over all walkers, in walker blocks
// PHMSD::energy_alg0 (PHMSD.cpp:122-370), COLLINEAR walkers, in walker blocks: each unique
// configuration's own G and single-determinant energy, its overlap relative to the spin's
// reference configuration, the up spin's Cholesky traces carried to the down configurations
// through the couplings (KEr[d] = sum_u C[d][u] ratio_up[u] KE_up[u]), then the CI-weighted sums.
void ph_energy_alg0(afq_ph_t* ph, hipStream_t s, int nw, const double* d_wbuf, long ldw,
                    double* d_E, double* d_ov) {
  const int nmo = ph->nmo, na = ph->nact, K = ph->nchol;
  Spin* S = ph->spin;
  const int nu = S[0].nunique, nd = S[1].nunique;
  const long em1 = 8L * std::max(nu, nd) * ph->wblock;   // doubles per spin of Emsd
  for (int w0 = 0; w0 < nw; w0 += ph->wblock) {
    const int nb = std::min(ph->wblock, nw - w0);
    const double* Phi[2] = {d_wbuf + 2 * (long)w0 * ldw,
                            d_wbuf + 2 * ((long)w0 * ldw + (long)nmo * S[0].nel)};
    launch_ph_fill(s, nb, 0.0, 0.0, ph->log_ov);
    for (int sp = 0; sp < 2; ++sp) {
      launch_ph0_products(s, nb, nmo, S[sp].nel, na, K, Phi[sp], ldw, S[sp].A, S[sp].haj,
                          S[sp].Lna, ph->G[sp], ph->QQ0[sp], ph->Swia, ph->Twina);
      launch_ph0_config(s, nb, S[sp].nunique, S[sp].nel, na, K, S[sp].dconf, ph->QQ0[sp],
                        ph->Swia, ph->Twina, sp == 0 ? ph->E0 : 0.0, ph->ratios[sp],
                        ph->Emsd + sp * em1, sp == 0 ? ph->KEa : nullptr,
                        sp == 1 ? ph->KEb : nullptr, ph->Tinv0);
      launch_ph0_ratios(s, nb, S[sp].nunique, ph->ratios[sp], ph->log_ov);
      if (sp == 0)
        launch_ph_couple_ke(s, nb, K, nd, S[1].cp, S[1].ci, S[1].cv, ph->ratios[0], ph->KEa,
                            nullptr, ph->KEb);
    }
    launch_ph_spin_weights(s, nb, nu, S[0].cp, S[0].ci, S[0].cv, ph->ratios[0], ph->ratios[1],
                           ph->wgt[0], ph->Ov);
    launch_ph_spin_weights(s, nb, nd, S[1].cp, S[1].ci, S[1].cv, ph->ratios[1], ph->ratios[0],
                           ph->wgt[1], ph->Ov2);
    launch_ph0_assemble(s, nb, nu, nd, ph->wgt[0], ph->Emsd, ph->wgt[1], ph->Emsd + em1,
                        ph->ratios[1], ph->acc);
    launch_ph_finalize(s, nb, ph->acc, ph->Ov, ph->log_ov, d_E + 6L * w0, d_ov + 2L * w0);
  }
}

void ph_energy_all(afq_ph_t* ph, hipStream_t s, int nw, const double* d_wbuf, long ldw,
                   double* d_E, double* d_ov) {
    if (ph->energy_alg == AFQ_PH_ENERGY_REFERENCE) {
      ph_energy_alg0(ph, s, nw, d_wbuf, ldw, d_E, d_ov);
      return;
    }
    const int nmo = ph->nmo, na = ph->nact, K = ph->nchol;
    Spin* S = ph->spin;
    const int nu = S[0].nunique, nd = S[1].nunique;
    for (int w0 = 0; w0 < nw; w0 += ph->wblock) {
      const int nb = std::min(ph->wblock, nw - w0);
      const double* Phi[2] = {d_wbuf + 2 * (long)w0 * ldw,
                              d_wbuf + 2 * ((long)w0 * ldw + (long)nmo * S[0].nel)};
      launch_ph_fill(s, 3L * nb, 0.0, 0.0, ph->acc);
      // 1. reference G, log overlap, QQ0 and the excitation overlap ratios of both spins
      ph_overlaps_block(ph, s, nb, Phi, ldw);
      // 2a. weights and the multi-determinant overlap
      launch_ph_spin_weights(s, nb, nu, S[0].cp, S[0].ci, S[0].cv, ph->ratios[0], ph->ratios[1],
                             ph->wgt[0], ph->Ov);
      // 3a. reference energy of the up spin (KE row 0 = its Cholesky trace), times Ov
      launch_ph_reference_energy(s, nb, S[0].nel, nmo, K, na, 1.0, ph->E0, ph->G[0], S[0].Lna,
                                 S[0].haj, ph->Eref, ph->KEa, ph->Twina, ph->Swia);
      launch_ph_accumulate(s, nb, ph->Eref, ph->Ov, nullptr, ph->acc);
      // 4a. excited up determinants: E corrections and their Cholesky vectors KE (rows 1..)
      launch_ph_fill(s, 3L * nb, 0.0, 0.0, ph->Ex);
      for (const Shell& sh : S[0].shells)
        for (int d0 = 0; d0 < sh.ndet; d0 += kPhDetBlock) {
          const int nbk = std::min(kPhDetBlock, sh.ndet - d0);
          const int* iex = sh.iexcit + 2L * sh.nex * d0;
          const double* w = ph->wgt[0] + 2L * (sh.row + d0) * nb;
          launch_ph_compact_R(s, nb, nbk, sh.nex, na, S[0].nel, iex, S[0].refc, ph->QQ0[0],
                              ph->R);
          launch_ph_1body(s, nb, nbk, sh.nex, na, S[0].nel, iex, S[0].refc, ph->Swia, ph->R, w,
                          ph->Ex);
          launch_ph_2body(s, nb, nbk, sh.nex, na, S[0].nel, K, iex, S[0].refc, ph->Twina, ph->R,
                          w, ph->Ex + 2L * nb, ph->Ex + 4L * nb,
                          ph->KEa + 2L * (sh.row + d0) * nb * K, ph->p2ws, ph->p2ws_doubles);
        }
      launch_ph_accumulate(s, nb, nullptr, nullptr, ph->Ex, ph->acc);
      // KEb[d] = ratio_dn[d] sum_u C[d][u] ratio_up[u] KEa[u]
      launch_ph_couple_ke(s, nb, K, nd, S[1].cp, S[1].ci, S[1].cv, ph->ratios[0], ph->KEa,
                          ph->ratios[1], ph->KEb);
      // 2b. weights of the down determinants
      launch_ph_spin_weights(s, nb, nd, S[1].cp, S[1].ci, S[1].cv, ph->ratios[1], ph->ratios[0],
                             ph->wgt[1], ph->Ov2);
      // 3b. reference energy of the down spin (E0 only with the up spin), the cross term
      launch_ph_reference_energy(s, nb, S[1].nel, nmo, K, na, 1.0, 0.0, ph->G[1], S[1].Lna,
                                 S[1].haj, ph->Eref, ph->KEa, ph->Twina, ph->Swia);
      launch_ph_accumulate(s, nb, ph->Eref, ph->Ov, nullptr, ph->acc);
      launch_ph_ke_dot(s, nb, K, 1, ph->KEb, ph->KEa, ph->acc + 4, 3);
      // 4b. excited down determinants, contracted with the coupled up vectors
      launch_ph_fill(s, 3L * nb, 0.0, 0.0, ph->Ex);
      for (const Shell& sh : S[1].shells)
        for (int d0 = 0; d0 < sh.ndet; d0 += kPhDetBlock) {
          const int nbk = std::min(kPhDetBlock, sh.ndet - d0);
          const int* iex = sh.iexcit + 2L * sh.nex * d0;
          const double* w = ph->wgt[1] + 2L * (sh.row + d0) * nb;
          launch_ph_compact_R(s, nb, nbk, sh.nex, na, S[1].nel, iex, S[1].refc, ph->QQ0[1],
                              ph->R);
          launch_ph_1body(s, nb, nbk, sh.nex, na, S[1].nel, iex, S[1].refc, ph->Swia, ph->R, w,
                          ph->Ex);
          launch_ph_2body(s, nb, nbk, sh.nex, na, S[1].nel, K, iex, S[1].refc, ph->Twina, ph->R,
                          w, ph->Ex + 2L * nb, ph->Ex + 4L * nb, ph->KEl, ph->p2ws,
                          ph->p2ws_doubles);
          launch_ph_ke_dot(s, nb, K, nbk, ph->KEb + 2L * (sh.row + d0) * nb * K, ph->KEl,
                           ph->Ex + 4L * nb, 1);
        }
      launch_ph_accumulate(s, nb, nullptr, nullptr, ph->Ex, ph->acc);
      // E /= Ov, Ov -> log Ov + log of the reference overlaps
      launch_ph_finalize(s, nb, ph->acc, ph->Ov, ph->log_ov, d_E + 6L * w0, d_ov + 2L * w0);
    }
}

// PHMSD::vbias over all walkers, in walker blocks
void ph_vbias_all(afq_ph_t* ph, hipStream_t s, int nw, const double* d_wbuf, long ldw, double dt,
                  double* d_v) {
    const int nmo = ph->nmo, na = ph->nact, K = ph->nchol;
    const int kx = 2 * na * nmo;
    for (int w0 = 0; w0 < nw; w0 += ph->wblock) {
      const int nb = std::min(ph->wblock, nw - w0);
      const double* Phi[2] = {d_wbuf + 2 * (long)w0 * ldw,
                              d_wbuf + 2 * ((long)w0 * ldw + (long)nmo * ph->spin[0].nel)};
      ph_density_block(ph, s, nb, Phi, ldw, ph->Gblk, ph->Eref);   // log overlap unused
      launch_gemm_nt_z(s, AFQ_FP64, nb, K, kx, std::sqrt(dt), ph->Gblk, kx, ph->Lstack, kx, true,
                       d_v + 2L * w0 * K, K, ph->split, ph->split_elems);
    }
}

// PHMSD::vMF (host density, device vbias); synchronises the stream
void ph_vmf(afq_ph_t* ph, hipStream_t st, double dt, double* d_vmf) {
    const int nmo = ph->nmo, na = ph->nact;
    // 1. overlap of every unique configuration relative to the reference: log det(As^H As),
    //    As the configuration's orbitals (columns of the conjugate-transposed orbital rows)
    std::vector<zc> ratio[2];
    for (int s = 0; s < 2; ++s) {
      const Spin& S = ph->spin[s];
      const int ne = S.nel;
      std::vector<zc> lds;
      for (const auto& c : S.confs) {
        std::vector<zc> T((size_t)ne * ne, 0.0);
        for (int i = 0; i < ne; ++i)
          for (int j = 0; j < ne; ++j) {
            zc acc = 0.0;
            for (int k = 0; k < nmo; ++k)
              acc += S.hA[(size_t)c[i] * nmo + k] * std::conj(S.hA[(size_t)c[j] * nmo + k]);
            T[(size_t)i * ne + j] = acc;
          }
        lds.push_back(host_logdet_inverse(T, ne, nullptr));
      }
      for (size_t d = 0; d < lds.size(); ++d) ratio[s].push_back(d ? std::exp(lds[d] - lds[0]) : 1.0);
    }
    zc ov = 0.0;
    {
      const Spin& S = ph->spin[0];
      for (int u = 0; u + 1 < (int)S.hcp.size(); ++u)
        for (int j = S.hcp[u]; j < S.hcp[u + 1]; ++j)
          ov += std::norm(S.hcv[j]) * ratio[0][u] * ratio[1][S.hci[j]];
    }
    // 2. diagonal and off-diagonal terms of each spin's density
    std::vector<zc> G((size_t)2 * na * nmo, 0.0);
    for (int s = 0; s < 2; ++s) {
      const int o = 1 - s;
      const Spin& S = ph->spin[s];
      const Spin& O = ph->spin[o];
      const int ne = S.nel;
      zc* Gs = G.data() + (size_t)s * na * nmo;
      for (size_t nd = 0; nd < S.confs.size(); ++nd) {
        const auto& c = S.confs[nd];
        // compact G of the configuration with itself: (As^H As)^-T As^T
        std::vector<zc> T((size_t)ne * ne), Ti;
        for (int i = 0; i < ne; ++i)
          for (int j = 0; j < ne; ++j) {
            zc acc = 0.0;
            for (int k = 0; k < nmo; ++k)
              acc += S.hA[(size_t)c[i] * nmo + k] * std::conj(S.hA[(size_t)c[j] * nmo + k]);
            T[(size_t)i * ne + j] = acc;
          }
        host_logdet_inverse(T, ne, &Ti);
        zc w = 0.0;
        for (int j = S.hcp[nd]; j < S.hcp[nd + 1]; ++j)
          w += ratio[o][S.hci[j]] * std::norm(S.hcv[j]);
        w *= ratio[s][nd];
        for (int i = 0; i < ne; ++i)
          for (int k = 0; k < nmo; ++k) {
            zc g = 0.0;
            for (int b = 0; b < ne; ++b)
              g += Ti[(size_t)b * ne + i] * std::conj(S.hA[(size_t)c[b] * nmo + k]);
            Gs[(size_t)c[i] * nmo + k] += w * g;
          }
      }
      // pairs of this spin's configurations in one row of the other spin's couplings that
      // differ by one orbital (no fermionic sign, the orbital rows as the reference adds them)
      for (int nd = 0; nd + 1 < (int)O.hcp.size(); ++nd)
        for (int n1 = O.hcp[nd]; n1 < O.hcp[nd + 1]; ++n1) {
          std::vector<int> c1 = S.confs[O.hci[n1]];
          std::sort(c1.begin(), c1.end());
          for (int n2 = n1 + 1; n2 < O.hcp[nd + 1]; ++n2) {
            std::vector<int> c2 = S.confs[O.hci[n2]];
            std::sort(c2.begin(), c2.end());
            int nh = 0, hole = -1, part = -1;
            for (int x : c1)
              if (!std::binary_search(c2.begin(), c2.end(), x)) {
                ++nh;
                hole = x;
              }
            if (nh != 1) continue;
            for (int x : c2)
              if (!std::binary_search(c1.begin(), c1.end(), x)) part = x;
            const zc w = O.hcv[n1] * std::conj(O.hcv[n2]);
            for (int k = 0; k < nmo; ++k) {
              Gs[(size_t)part * nmo + k] += w * S.hA[(size_t)hole * nmo + k];
              Gs[(size_t)hole * nmo + k] += std::conj(w) * S.hA[(size_t)part * nmo + k];
            }
          }
        }
    }
    for (auto& g : G) g /= ov;
    // 3. vbias of the mean-field density: sqrt(dt) sum_s sum_{a,k} Lna_s[n][a][k] G_s[a][k]
    std::vector<double> flat(2 * G.size());
    for (size_t e = 0; e < G.size(); ++e) {
      flat[2 * e] = G[e].real();
      flat[2 * e + 1] = G[e].imag();
    }
    AFQ_HIP(hipMemcpyAsync(ph->Gblk, flat.data(), flat.size() * sizeof(double),
                           hipMemcpyHostToDevice, st));
    const int kx = 2 * na * nmo;
    launch_gemm_nt_z(st, AFQ_FP64, 1, ph->nchol, kx, std::sqrt(dt), ph->Gblk, kx, ph->Lstack, kx,
                     true, d_vmf, ph->nchol);
    AFQ_HIP(hipStreamSynchronize(st));
}

// PHMSD overlap (the log overlap of MixedDensityMatrix without G) over all walkers
void ph_overlap_all(afq_ph_t* ph, hipStream_t s, int nw, const double* d_wbuf, long ldw,
                    double* d_ov) {
  const int nmo = ph->nmo;
  Spin* S = ph->spin;
  for (int w0 = 0; w0 < nw; w0 += ph->wblock) {
    const int nb = std::min(ph->wblock, nw - w0);
    const double* Phi[2] = {d_wbuf + 2 * (long)w0 * ldw,
                            d_wbuf + 2 * ((long)w0 * ldw + (long)nmo * S[0].nel)};
    ph_overlaps_block(ph, s, nb, Phi, ldw);
    launch_ph_spin_weights(s, nb, S[0].nunique, S[0].cp, S[0].ci, S[0].cv, ph->ratios[0],
                           ph->ratios[1], ph->wgt[0], ph->Ov);
    launch_ph_log_overlap(s, nb, ph->Ov, ph->log_ov, d_ov + 2L * w0);
  }
}

long ph_poff(const afq_ph_t* ph) { return (long)ph->nmo * (ph->spin[0].nel + ph->spin[1].nel); }

void ph_check_walkers(const afq_ph_t* ph, int W, long ldw, bool prepared) {
  AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "PHMSD: null handle");
  AFQ_CHECK(W >= 0, AFQ_EBADSHAPE, "PHMSD: negative walker count");
  AFQ_CHECK(ldw >= ph_poff(ph) + AFQ_P_COUNT, AFQ_EBADSHAPE,
            "PHMSD: ldw smaller than the Slater matrices and the walker properties");
  if (prepared) {
    AFQ_CHECK(ph->dt > 0.0, AFQ_EINVAL, "PHMSD: afq_ph_prepare has not been called");
    AFQ_CHECK(W <= ph->wcap, AFQ_ECAPACITY, "PHMSD: more walkers than afq_ph_prepare sized for");
  }
}

}  // namespace

extern "C" {

int afq_ph_create(int nmo, int nup, int ndn, int nchol, int ndet, const int* occs,
                  const double* ci, int nref, const double* orbmats, const double* h1,
                  const double* likn, double E0, int walker_block, afq_ph* out) {
  return ph_guard([&] {
    AFQ_CHECK(out != nullptr, AFQ_EINVAL, "afq_ph_create: null output handle");
    *out = nullptr;
    AFQ_CHECK(nmo > 0 && nup > 0 && ndn > 0 && nup <= nmo && ndn <= nmo && nchol > 0 &&
                  ndet > 0 && walker_block > 0,
              AFQ_EBADSHAPE, "afq_ph_create: bad sizes");
    AFQ_CHECK(nmo <= 256 && nup <= 64 && ndn <= 64, AFQ_EUNSUPPORTED,
              "afq_ph_create: needs nmo <= 256 and at most 64 electrons per spin");
    AFQ_CHECK(nref == 1 || nref == 2, AFQ_EBADSHAPE, "afq_ph_create: nref must be 1 or 2");
    AFQ_CHECK(occs && ci && orbmats && h1 && likn, AFQ_EINVAL, "afq_ph_create: null input");
    const int nel[2] = {nup, ndn};
    // --- build_ph_struct on the host
    std::vector<int> ref[2];
    for (int s = 0; s < 2; ++s) {
      for (int k = 0; k < nel[s]; ++k) {
        const int q = occs[(s ? nup : 0) + k];
        AFQ_CHECK(q >= s * nmo && q < (s + 1) * nmo, AFQ_EINVAL,
                  "afq_ph_create: bad occupation in the reference determinant");
        ref[s].push_back(q);
      }
      AFQ_CHECK(std::is_sorted(ref[s].begin(), ref[s].end()), AFQ_EINVAL,
                "afq_ph_create: occupations must be sorted");
    }
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < nel[s]; ++k)
        AFQ_CHECK(ref[s][k] - s * nmo == k, AFQ_EUNSUPPORTED,
                  "afq_ph_create: the reference must occupy orbitals 0..nel-1 of each spin (as "
                  "the reference's energy assumes)");
    std::vector<std::vector<std::vector<int>>> unique[2];   // [spin][rank] list of excitations
    for (int s = 0; s < 2; ++s) unique[s].resize(nel[s] + 1);
    struct Conf {
      int rank[2], loc[2];
      double re, im;
    };
    std::vector<Conf> confs((size_t)ndet);
    std::vector<int> conf, ex;
    for (int d = 0; d < ndet; ++d) {
      Conf c{};
      c.re = ci[2L * d];
      c.im = ci[2L * d + 1];
      for (int s = 0; s < 2; ++s) {
        conf.assign(occs + (long)d * (nup + ndn) + (s ? nup : 0),
                    occs + (long)d * (nup + ndn) + (s ? nup : 0) + nel[s]);
        for (int q : conf)
          AFQ_CHECK(q >= s * nmo && q < (s + 1) * nmo, AFQ_EINVAL,
                    "afq_ph_create: bad occupation number in determinant " + std::to_string(d));
        AFQ_CHECK(std::is_sorted(conf.begin(), conf.end()), AFQ_EINVAL,
                  "afq_ph_create: occupations must be sorted");
        const int sign = excitation_of(ref[s], conf, ex);
        if (sign < 0) {
          c.re = -c.re;
          c.im = -c.im;
        }
        const int n = (int)ex.size() / 2;
        for (int k = n; k < 2 * n; ++k) ex[k] -= s * nmo;      // active indices
        c.rank[s] = n;
        c.loc[s] = 0;
        if (n > 0) {
          auto& lst = unique[s][n];
          auto it = std::find(lst.begin(), lst.end(), ex);
          c.loc[s] = (int)(it - lst.begin());
          if (it == lst.end()) lst.push_back(ex);
        }
      }
      confs[d] = c;
    }
    std::unique_ptr<afq_ph_t> P(new afq_ph_t);
    AFQ_HIP(hipGetDevice(&P->device));
    P->nmo = nmo;
    P->nact = nmo;
    P->nchol = nchol;
    P->E0 = E0;
    P->wblock = walker_block;
    auto& own = P->owned;
    std::vector<int> first[2];
    int max_nex = 1;
    for (int s = 0; s < 2; ++s) {
      Spin& S = P->spin[s];
      S.nel = nel[s];
      first[s].assign(nel[s] + 1, 0);
      int row = 1;
      for (int n = 1; n <= nel[s]; ++n) {
        first[s][n] = row;
        const auto& lst = unique[s][n];
        if (lst.empty()) continue;
        std::vector<int> flat;
        for (const auto& e : lst) flat.insert(flat.end(), e.begin(), e.end());
        Shell sh;
        sh.nex = n;
        sh.ndet = (int)lst.size();
        sh.row = row;
        sh.iexcit = dev_copy(own, flat.data(), flat.size());
        S.shells.push_back(sh);
        row += sh.ndet;
        max_nex = std::max(max_nex, n);
      }
      S.nunique = row;
      S.confs.assign(1, std::vector<int>());
      for (int k = 0; k < nel[s]; ++k) S.confs[0].push_back(ref[s][k] - s * nmo);
      for (int n = 1; n <= nel[s]; ++n)
        for (const auto& e : unique[s][n]) {
          std::vector<int> c = S.confs[0];
          for (int p = 0; p < n; ++p) c[e[p]] = e[n + p];
          S.confs.push_back(c);
        }
      std::vector<int> rc(nel[s]);
      for (int k = 0; k < nel[s]; ++k) rc[k] = ref[s][k] - s * nmo;
      S.refc = dev_copy(own, rc.data(), rc.size());
      std::vector<int> flat;
      for (const auto& c : S.confs) flat.insert(flat.end(), c.begin(), c.end());
      S.dconf = dev_copy(own, flat.data(), flat.size());
    }
    AFQ_CHECK(max_nex <= 8, AFQ_EUNSUPPORTED, "afq_ph_create: excitations of rank > 8");
    // couplings: spin 0 rows u, columns d; spin 1 rows d, columns u; entries sorted by column
    for (int s = 0; s < 2; ++s) {
      std::vector<std::vector<std::pair<int, std::pair<double, double>>>> rows(
          (size_t)P->spin[s].nunique);
      for (const Conf& c : confs) {
        const int u = c.rank[0] ? first[0][c.rank[0]] + c.loc[0] : 0;
        const int d = c.rank[1] ? first[1][c.rank[1]] + c.loc[1] : 0;
        const int r = s == 0 ? u : d, col = s == 0 ? d : u;
        rows[r].push_back({col, {c.re, -c.im}});
      }
      std::vector<int> cp(1, 0), cix;
      std::vector<double> cv;
      for (auto& r : rows) {
        std::stable_sort(r.begin(), r.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        for (const auto& e : r) {
          cix.push_back(e.first);
          cv.push_back(e.second.first);
          cv.push_back(e.second.second);
        }
        cp.push_back((int)cix.size());
      }
      P->spin[s].hcp = cp;
      P->spin[s].hci = cix;
      for (size_t e = 0; e < cix.size(); ++e)
        P->spin[s].hcv.emplace_back(cv[2 * e], cv[2 * e + 1]);
      P->spin[s].cp = dev_copy(own, cp.data(), cp.size());
      P->spin[s].ci = dev_copy(own, cix.data(), cix.size());
      P->spin[s].cv = dev_copy(own, cv.data(), cv.size());
    }
    // --- half-rotated Hamiltonian per spin on the GPU
    hipStream_t st = nullptr;
    const size_t nmm = (size_t)nmo * nmo;
    double* dh = dev_copy(own, h1, nmm);
    double* dL = dev_copy(own, likn, nmm * nchol);
    P->h = dh;
    P->Likn = dL;
    for (int s = 0; s < 2; ++s) {
      Spin& S = P->spin[s];
      S.A = dev_copy(own, orbmats + (nref == 2 ? s : 0) * 2 * nmm, 2 * nmm);
      const double* ho = orbmats + (nref == 2 ? s : 0) * 2 * nmm;
      S.hA.resize(nmm);
      for (size_t e = 0; e < nmm; ++e) S.hA[e] = {ho[2 * e], ho[2 * e + 1]};
      S.Lna = dev_alloc<double>(own, 2 * nmm * nchol);
      S.haj = dev_alloc<double>(own, 2 * nmm);
      launch_half_rotate(st, nmo, nmo, nchol, S.A, dh, dL, S.haj, S.Lna);
    }
    {
      double* stackA = dev_alloc<double>(own, 4 * nmm);
      AFQ_HIP(hipMemcpy(stackA, P->spin[0].A, 2 * nmm * sizeof(double), hipMemcpyDeviceToDevice));
      AFQ_HIP(hipMemcpy(stackA + 2 * nmm, P->spin[1].A, 2 * nmm * sizeof(double),
                        hipMemcpyDeviceToDevice));
      P->Lstack = dev_alloc<double>(own, 4 * nmm * nchol);
      double* hstack = dev_alloc<double>(own, 4 * nmm);
      launch_half_rotate(st, nmo, 2 * nmo, nchol, stackA, dh, dL, hstack, P->Lstack);
    }
    AFQ_HIP(hipStreamSynchronize(st));
    // --- workspace for one walker block
    const long nb = walker_block, K = nchol, na = nmo;
    const int maxu = std::max(P->spin[0].nunique, P->spin[1].nunique);
    int maxdet = 1;
    for (int s = 0; s < 2; ++s)
      for (const Shell& sh : P->spin[s].shells) maxdet = std::max(maxdet, sh.ndet);
    const long dblk = std::min(maxdet, kPhDetBlock);
    const int maxel = std::max(nup, ndn);
    for (int s = 0; s < 2; ++s) {
      P->G[s] = dev_alloc<double>(own, 2 * nb * nel[s] * nmo);
      P->QQ0[s] = dev_alloc<double>(own, 2 * nb * na * nel[s]);
      P->ratios[s] = dev_alloc<double>(own, 2 * nb * P->spin[s].nunique);
      P->wgt[s] = dev_alloc<double>(own, 2 * nb * P->spin[s].nunique);
    }
    P->work = dev_alloc<double>(own, 2L * maxel * nmo);
    P->Ov = dev_alloc<double>(own, 2 * nb);
    P->Ov2 = dev_alloc<double>(own, 2 * nb);
    P->log_ov = dev_alloc<double>(own, 2 * nb);
    P->Eref = dev_alloc<double>(own, 6 * nb);
    P->Ex = dev_alloc<double>(own, 6 * nb);
    P->acc = dev_alloc<double>(own, 6 * nb);
    P->Twina = dev_alloc<double>(own, 2 * nb * maxel * K * na);
    P->Swia = dev_alloc<double>(own, 2 * nb * maxel * na);
    P->KEa = dev_alloc<double>(own, 2L * maxu * nb * K);
    P->KEb = dev_alloc<double>(own, 2L * maxu * nb * K);
    P->KEl = dev_alloc<double>(own, 2 * std::max(dblk, 1L) * nb * K);
    P->R = dev_alloc<double>(own, 2 * nb * dblk * max_nex * na);
    P->Racc = dev_alloc<double>(own, 2 * nb * maxel * na);
    P->Gblk = dev_alloc<double>(own, 2 * nb * 2 * na * nmo);
    P->Emsd = dev_alloc<double>(own, 16L * maxu * nb);
    if (maxel <= 32) P->Tinv0 = dev_alloc<double>(own, 2L * maxu * nb * maxel * maxel);
    P->p2ws_doubles = ph_2body_workspace((int)nb, (int)na, maxel, (int)K);
    if (P->p2ws_doubles > 0) P->p2ws = dev_alloc<double>(own, P->p2ws_doubles);
    P->split_elems = (long)kMaxSplits * nb * K;
    P->split = dev_alloc<double>(own, 2 * P->split_elems);
    *out = P.release();
  });
}

void afq_ph_destroy(afq_ph ph) { delete ph; }

int afq_ph_info(afq_ph ph, int* nunique, int* nshells) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "afq_ph_info: null handle");
    for (int s = 0; s < 2; ++s) {
      if (nunique) nunique[s] = ph->spin[s].nunique;
      if (nshells) nshells[s] = (int)ph->spin[s].shells.size();
    }
  });
}

int afq_ph_set_energy_algorithm(afq_ph ph, int algorithm) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "afq_ph_set_energy_algorithm: null handle");
    AFQ_CHECK(algorithm == AFQ_PH_ENERGY_REFERENCE || algorithm == AFQ_PH_ENERGY_WOODBURY,
              AFQ_EINVAL, "afq_ph_set_energy_algorithm: unknown algorithm");
    if (algorithm == AFQ_PH_ENERGY_REFERENCE)
      AFQ_CHECK(ph->spin[0].nel <= 32 && ph->spin[1].nel <= 32, AFQ_EUNSUPPORTED,
                "afq_ph_set_energy_algorithm: the reference algorithm takes at most 32 "
                "electrons per spin");
    ph->energy_alg = algorithm;
  });
}

int afq_ph_energy(afq_ph ph, int nw, const double* d_wbuf, long ldw, double* d_E, double* d_ov,
                  void* stream) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "afq_ph_energy: null handle");
    AFQ_CHECK(nw >= 0, AFQ_EBADSHAPE, "afq_ph_energy: negative walker count");
    const long need = (long)ph->nmo * (ph->spin[0].nel + ph->spin[1].nel);
    AFQ_CHECK(ldw >= need, AFQ_EBADSHAPE, "afq_ph_energy: ldw smaller than the Slater matrices");
    if (nw == 0) return;
    AFQ_CHECK(d_wbuf && d_E && d_ov, AFQ_EINVAL, "afq_ph_energy: null operand");
    ph_energy_all(ph, reinterpret_cast<hipStream_t>(stream), nw, d_wbuf, ldw, d_E, d_ov);
  });
}

int afq_ph_mixed_density_matrix(afq_ph ph, int nw, const double* d_wbuf, long ldw, int full,
                                double* d_G, double* d_ov, void* stream) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "afq_ph_mixed_density_matrix: null handle");
    AFQ_CHECK(nw >= 0, AFQ_EBADSHAPE, "afq_ph_mixed_density_matrix: negative walker count");
    const long need = (long)ph->nmo * (ph->spin[0].nel + ph->spin[1].nel);
    AFQ_CHECK(ldw >= need, AFQ_EBADSHAPE,
              "afq_ph_mixed_density_matrix: ldw smaller than the Slater matrices");
    if (nw == 0) return;
    AFQ_CHECK(d_wbuf && d_G && d_ov, AFQ_EINVAL, "afq_ph_mixed_density_matrix: null operand");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int nmo = ph->nmo, na = ph->nact;
    for (int w0 = 0; w0 < nw; w0 += ph->wblock) {
      const int nb = std::min(ph->wblock, nw - w0);
      const double* Phi[2] = {d_wbuf + 2 * (long)w0 * ldw,
                              d_wbuf + 2 * ((long)w0 * ldw + (long)nmo * ph->spin[0].nel)};
      if (full) {
        ph_density_block(ph, s, nb, Phi, ldw, ph->Gblk, d_ov + 2L * w0);
        launch_ph_full_g(s, nb, na, nmo, ph->spin[0].A, ph->spin[1].A, ph->Gblk,
                         d_G + 2L * w0 * 2 * nmo * nmo);
      } else {
        ph_density_block(ph, s, nb, Phi, ldw, d_G + 2L * w0 * 2 * na * nmo, d_ov + 2L * w0);
      }
    }
  });
}

int afq_ph_vbias(afq_ph ph, int nw, const double* d_wbuf, long ldw, double dt, double* d_v,
                 void* stream) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "afq_ph_vbias: null handle");
    AFQ_CHECK(nw >= 0 && dt > 0.0, AFQ_EBADSHAPE, "afq_ph_vbias: bad walker count or dt");
    const long need = (long)ph->nmo * (ph->spin[0].nel + ph->spin[1].nel);
    AFQ_CHECK(ldw >= need, AFQ_EBADSHAPE, "afq_ph_vbias: ldw smaller than the Slater matrices");
    if (nw == 0) return;
    AFQ_CHECK(d_wbuf && d_v, AFQ_EINVAL, "afq_ph_vbias: null operand");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    ph_vbias_all(ph, s, nw, d_wbuf, ldw, dt, d_v);
  });
}

int afq_ph_vmf(afq_ph ph, double dt, double* d_vmf, void* stream) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "afq_ph_vmf: null handle");
    AFQ_CHECK(d_vmf != nullptr && dt > 0.0, AFQ_EINVAL, "afq_ph_vmf: null output or dt <= 0");
    ph_vmf(ph, reinterpret_cast<hipStream_t>(stream), dt, d_vmf);
  });
}

int afq_ph_prepare(afq_ph ph, double dt, int W, void* stream) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr, AFQ_EINVAL, "afq_ph_prepare: null handle");
    AFQ_CHECK(dt > 0.0 && W >= 0, AFQ_EINVAL, "afq_ph_prepare: dt <= 0 or negative W");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int M = ph->nmo, K = ph->nchol;
    auto& own = ph->step_owned;
    if (ph->vMF == nullptr) {
      ph->vMF = dev_alloc<double>(own, 2 * (size_t)K);
      ph->v0 = dev_alloc<double>(own, (size_t)M * M);
      ph->H1 = dev_alloc<double>(own, 2 * (size_t)M * M);
      ph->P1 = dev_alloc<double>(own, 2 * (size_t)M * M);
      ph->stats = dev_alloc<unsigned long long>(own, 7);
      AFQ_HIP(hipMemset(ph->stats, 0, 7 * sizeof(unsigned long long)));
      afq_default_params(&ph->params);
    }
    // the mean field of the trial with its imaginary part dropped (AFQMCBasePropagator.cpp:
    // 85-107), v0, and the one-body propagator built from them
    ph_vmf(ph, s, dt, ph->vMF);
    launch_zero_imag(s, K, ph->vMF);
    launch_v0(s, M, K, ph->Likn, ph->v0);
    build_one_body_propagator(s, M, K, dt, ph->h, ph->v0, ph->Likn, ph->vMF, ph->H1, ph->P1);
    ph->dt = dt;
    if (W > ph->wcap) {
      const size_t w = (size_t)W;
      ph->X = dev_alloc<double>(own, 2 * w * K);
      ph->HW = dev_alloc<double>(own, 2 * w);
      ph->mf = dev_alloc<double>(own, 2 * w);
      ph->V = dev_alloc<double>(own, 2 * w * M * M);
      ph->ovw = dev_alloc<double>(own, 2 * w);
      ph->Ew = dev_alloc<double>(own, 6 * w);
      ph->cum = dev_alloc<double>(own, w);
      ph->off = dev_alloc<int>(own, w);
      ph->src = dev_alloc<int>(own, w);
      ph->ok = dev_alloc<int>(own, 1);
      ph->wcap = W;
    }
  });
}

int afq_ph_set_params(afq_ph ph, const afq_step_params* p) {
  return ph_guard([&] {
    AFQ_CHECK(ph != nullptr && p != nullptr, AFQ_EINVAL, "afq_ph_set_params: null argument");
    AFQ_CHECK(ph->dt > 0.0, AFQ_EINVAL, "afq_ph_set_params: call afq_ph_prepare first");
    ph->params = *p;
  });
}

int afq_ph_seed_properties(afq_ph ph, int W, double* d_wbuf, long ldw, void* stream) {
  return ph_guard([&] {
    ph_check_walkers(ph, W, ldw, true);
    if (W == 0) return;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    ph_energy_all(ph, s, W, d_wbuf, ldw, ph->Ew, ph->ovw);
    launch_set_properties(s, W, d_wbuf, ldw, ph_poff(ph), ph->ovw, ph->Ew, 1);
  });
}

int afq_ph_step(afq_ph ph, int W, double* d_wbuf, long ldw, const double* d_xi, double Eshift,
                void* stream) {
  return ph_guard([&] {
    ph_check_walkers(ph, W, ldw, true);
    if (W == 0) return;
    AFQ_CHECK(d_wbuf && d_xi, AFQ_EINVAL, "afq_ph_step: null operand");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int M = ph->nmo, K = ph->nchol;
    const afq_step_params& p = ph->params;
    // 1. vbias of the multi-determinant G (PHMSD::vbias), 2. X, HW, mean-field factor
    ph_vbias_all(ph, s, W, d_wbuf, ldw, ph->dt, ph->X);
    launch_construct_X(s, W, K, p.vbias_bound, p.use_cp_constraint || p.project_force_bias,
                       p.free_projection, ph->vMF, d_xi, ph->X, ph->HW, ph->mf);
    // 3. vHS = sqrt(dt) L X (Real3IndexFactorization::vHS), 4. both spins' Slater matrices
    launch_gemm_nt_z(s, AFQ_FP64, W, M * M, K, std::sqrt(ph->dt), ph->X, K, ph->Likn, K, false,
                     ph->V, (long)M * M);
    launch_propagate(s, W, M, ph->spin[0].nel, d_wbuf, ldw, ph->P1, ph->V, p.taylor_n);
    launch_propagate(s, W, M, ph->spin[1].nel, d_wbuf + 2L * M * ph->spin[0].nel, ldw, ph->P1,
                     ph->V, p.taylor_n);
    // 5. new overlap (hybrid) or local energy and overlap (PHMSD::energy)
    if (p.hybrid)
      ph_overlap_all(ph, s, W, d_wbuf, ldw, ph->ovw);
    else
      ph_energy_all(ph, s, W, d_wbuf, ldw, ph->Ew, ph->ovw);
    // 6.-7. weights, phases, energies, bounds
    launch_walker_update(s, W, d_wbuf, ldw, ph_poff(ph), p, ph->dt, Eshift, ph->E0, ph->ovw,
                         ph->Ew, ph->mf, ph->HW, p.hybrid, ph->stats);
  });
}

int afq_ph_orthogonalize(afq_ph ph, int W, double* d_wbuf, long ldw, void* stream) {
  return ph_guard([&] {
    ph_check_walkers(ph, W, ldw, true);
    if (W == 0) return;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int M = ph->nmo;
    launch_orthogonalize(s, W, M, ph->spin[0].nel, d_wbuf, ldw, nullptr);
    launch_orthogonalize(s, W, M, ph->spin[1].nel, d_wbuf + 2L * M * ph->spin[0].nel, ldw,
                         nullptr);
    ph_overlap_all(ph, s, W, d_wbuf, ldw, ph->ovw);
    launch_set_properties(s, W, d_wbuf, ldw, ph_poff(ph), ph->ovw, nullptr, 0);
  });
}

int afq_ph_population_control(afq_ph ph, int W, double* d_wbuf, long ldw, double uniform,
                              void* stream) {
  return ph_guard([&] {
    ph_check_walkers(ph, W, ldw, true);
    AFQ_CHECK(uniform >= 0.0 && uniform < 1.0, AFQ_EINVAL, "uniform must be in [0, 1)");
    if (W == 0) return;
    launch_population_control(reinterpret_cast<hipStream_t>(stream), W, d_wbuf, ldw,
                              ph_poff(ph), uniform, ph->cum, ph->off, ph->src, ph->ok, ph->stats);
  });
}

}  // extern "C"
