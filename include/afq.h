/* afq.h -- C ABI of the MI355X (gfx950) phaseless-AFQMC walker step.
 *
 * Drop-in for SAFIRE's walker step with a Cholesky-factorised Hamiltonian: the operations of
 * HamiltonianOperations / Real3IndexFactorization (vbias, vHS, energy), of det_ops
 * (MixedDensityMatrix, Log_Overlap, Propagate, orthogonalize) and of AFQMCBasePropagator
 * (assemble_X, walker_update, Propagate, Orthogonalize), acting on a WalkerSet buffer.
 *
 * Conventions (identical to the reference so that results diff elementwise):
 *   - C order, complex<double> interleaved (re, im): "c128".
 *   - Walker buffer: c128 [W][ldw]; walker w's Slater matrix Phi_w[i][a] (i < M orbital,
 *     a < N electron) at w*ldw + i*N + a, then the 11 property slots at w*ldw + M*N + p
 *     (p = AFQ_P_WEIGHT ... AFQ_P_THETA) -- WalkerSetBase.cpp:43-124, WalkerSetBase.h:255-328.
 *   - Compact mixed density matrix G: c128 [W][N][M] (= (Phi T^-1)^T, density_matrix.hpp:574).
 *   - vbias / X: c128 [W][K];  vHS V: c128 [W][M][M];  E: c128 [W][3] = (E1 incl. E0, EXX, EJ).
 *   - Pointers named d_* are device pointers; h_* are host pointers. `stream` is a hipStream_t
 *     (NULL = the default stream); every call only enqueues work on it.
 *   - Every entry point returns an afq_status; nothing throws across the ABI. afq_last_error()
 *     returns the message of the last failure on the calling thread.
 *   - Nothing allocates device memory after afq_ham_reserve(), so afq_step() and the other
 *     per-step calls may be captured in a hipGraph.
 */
#ifndef SAFIRE_AMD_AFQ_H
#define SAFIRE_AMD_AFQ_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afq_ham_s* afq_ham;
typedef struct afq_rng_t* afq_rng;
typedef struct afq_ph_t* afq_ph;

enum afq_status {
  AFQ_OK = 0,
  AFQ_EBADSHAPE = 1,   /* utils::check shape failures of the reference */
  AFQ_EHIP = 2,        /* a HIP runtime call failed */
  AFQ_ENOMEM = 3,
  AFQ_EINVAL = 4,
  AFQ_EUNSUPPORTED = 5,
  AFQ_ECAPACITY = 6    /* more walkers than afq_ham_reserve() sized the workspace for */
};

/* WALKER_TYPES values of the reference (config.h:47-51). */
enum afq_walker_type { AFQ_CLOSED = 1, AFQ_COLLINEAR = 2 };

/* Arithmetic of the Cholesky contractions (vbias, vHS, exchange/Coulomb energy). Overlaps,
 * log-determinants, inverses, QR, the propagator, weights and energy sums stay FP64. */
enum afq_precision {
  AFQ_FP64 = 0,        /* plain form: everything FP64 (FP64 MFMA)                          */
  AFQ_FP32 = 1,        /* contractions in FP32 on MFMA with FP32 accumulation               */
  AFQ_BF16X3 = 2,      /* mixed form: the exchange energy on BF16 MFMA with both operands
                          split into three BF16 slices (hi/mid/lo, ~24 bits) and the six
                          slice products of order <= 2 accumulated in FP32, per real/imag
                          part; vbias and vHS as in AFQ_FP32. Needs N <= 16, M <= 96 and
                          M % 8 == 0 (AFQ_EUNSUPPORTED otherwise)                           */
  AFQ_INT8_SPLIT = 3   /* mixed form: the exchange energy on int8 MFMA with every row of G
                          and Lnak scaled by a power of two and split into 5 int8 slices
                          (x = u (q0 + q1/254 + ... + q4/254^4)); the 15 slice products of
                          order <= 4 accumulate exactly in int32 and are combined in FP64
                          (energy error ~1e-3 of AFQ_FP32's). vbias and vHS as in AFQ_FP32.
                          Needs N <= 32 and M <= 128 (AFQ_EUNSUPPORTED otherwise)           */
};

/* Walker property slots (WalkerConfig.hpp:28-48, offsets WalkerSetBase.cpp:87-117). */
enum afq_property {
  AFQ_P_WEIGHT = 0, AFQ_P_PHASE, AFQ_P_PSEUDO_ELOC, AFQ_P_E1, AFQ_P_EXX, AFQ_P_EJ, AFQ_P_OVLP,
  AFQ_P_LOGSCL_UP, AFQ_P_LOGSCL_DN, AFQ_P_IS_UNITARY, AFQ_P_THETA, AFQ_P_COUNT
};

/* PropagatorParameters (parameters.hpp:103-126) with the defaults of resolve_defaults for a
 * non-model Hamiltonian (parameter_defaults.cpp:274-282). */
typedef struct {
  int taylor_n;                 /* 6 */
  double vbias_bound;           /* 50 */
  double upper_cutoff_scale;    /* 10 */
  double lower_cutoff_scale;    /* 1 */
  double weight_bound_floor;    /* 100 */
  double weight_bound_fraction; /* 0.1 */
  int hybrid;                   /* 1 */
  int free_projection;          /* 0 */
  int use_cp_constraint;        /* 0 */
  int project_force_bias;       /* 0 */
} afq_step_params;

void afq_default_params(afq_step_params* p);
const char* afq_last_error(void);
int afq_version(void);

/* ---------------------------------------------------------------------------------------------
 * Hamiltonian + trial + one-body propagator of one system.
 * Replaces RealDenseHamiltonian::getHamiltonianOperations (RealDenseHamiltonian.cpp:52-185:
 * haj, Lnak, v0), NOMSD construction (Wavefunction.cpp:256-287) and
 * AFQMCBasePropagator::generateP1 (AFQMCBasePropagator.cpp:42-214: vMF, H1, P1 = exp(-H1/2)).
 * Inputs are host arrays: h_h [M][M] f64 (hcore), h_Likn [M][M][K] f64 (DenseFactorized/L),
 * h_PsiT [N][M] c128 (trial orbitals as the conjugate transpose), one determinant with
 * coefficient 1, CLOSED walkers. All derived tensors are built on the GPU.
 * ------------------------------------------------------------------------------------------- */
int afq_ham_create(int M, int N, int K, const double* h_h, const double* h_Likn,
                   const double* h_PsiT, double E0, double dt, int walker_type, afq_ham* out);
int afq_ham_destroy(afq_ham ham);
/* Sizes the per-step workspace for up to max_walkers walkers (the only allocation after
 * create). Re-reserving with a larger count reallocates. */
int afq_ham_reserve(afq_ham ham, int max_walkers);
int afq_ham_set_precision(afq_ham ham, int precision);
int afq_ham_set_params(afq_ham ham, const afq_step_params* p);
/* Autotuner (no counterpart in the reference, which leaves tile shapes to its libraries): times
 * every tiling of the vbias and vHS GEMMs this precision has for W walkers (afq_ham_reserve
 * capacity) and keeps the fastest for the handle; results do not depend on the choice (every
 * tiling sums k in the same order). cache_path (NULL: none) is a JSON object of
 * "gemm/precision/W/shape": tiling entries, read first (a cached shape is not timed again) and
 * written back. cfg_out[2] (may be NULL) receives the (vbias, vHS) tilings. The walker buffer is
 * not touched; the workspace is. afq_ham_set_precision resets the tilings. */
int afq_autotune(afq_ham ham, int W, const char* cache_path, int* cfg_out, void* stream);
int afq_ham_get_params(afq_ham ham, afq_step_params* p);
/* Copies the derived tensors to host memory (any pointer may be NULL): haj [N][M] c128,
 * Lnak [K][N][M] c128, v0 [M][M] f64, vMF [K] c128, P1 [M][M] c128, H1 [M][M] c128. */
int afq_ham_tensors(afq_ham ham, double* h_haj, double* h_Lnak, double* h_v0, double* h_vMF,
                    double* h_P1, double* h_H1);
/* Local-energy bound statistics (total, upper, lower) and weight-bound statistics (total,
 * upper) since creation (AFQMCBasePropagator.h:199-237), and population-control (calls, calls
 * refused because the total weight collapsed below 1e-6). Any pointer may be NULL. */
int afq_ham_stats(afq_ham ham, long long* h_eloc3, long long* h_weight2, long long* h_popctl2);

/* ---------------------------------------------------------------------------------------------
 * det_ops level (density_matrix.hpp, propagate.hpp, orthogonalize.hpp): explicit operands.
 * ------------------------------------------------------------------------------------------- */
/* det_ops::MixedDensityMatrix (density_matrix.hpp:535-601), herm=true, compact:
 * T = PsiT Phi_w, d_ov[w] += wrapped log det T, d_G[w] = T^{-T} Phi_w^T ([N][M]); a singular T
 * gives G = 0 (getri_or_zero). d_G may be NULL (then this is det_ops::Log_Overlap,
 * density_matrix.hpp:437-447). d_PsiT [N][M] c128, d_Phi [nb][ldw] c128. */
int afq_det_mixed_density_matrix(int nb, int M, int N, const double* d_PsiT,
                                 const double* d_Phi, long ldw, double* d_G, double* d_ov,
                                 void* stream);
/* det_ops::detail::propagate_impl<'N'> (propagate.hpp:241-298) with a dense P1 [M][M] and
 * V [nb][M][M]: Phi <- P1 exp_taylor(iV, order) P1 Phi. */
int afq_det_propagate(int nb, int M, int N, double* d_Phi, long ldw, const double* d_P1,
                      const double* d_V, int order, void* stream);
/* det_ops::orthogonalize (orthogonalize.hpp:131-169): Phi <- Q diag(sign Re R_ii);
 * d_logdetR[w] += sum log(|R_ii|) (log_determinant_from_geqrf); d_logdetR may be NULL. */
int afq_det_orthogonalize(int nb, int M, int N, double* d_Phi, long ldw, double* d_logdetR,
                          void* stream);

/* ---------------------------------------------------------------------------------------------
 * Wavefunction / HamiltonianOperations level (NOMSD.cpp, Real3IndexFactorization.hpp).
 * ------------------------------------------------------------------------------------------- */
/* NOMSD::MixedDensityMatrix, CLOSED (NOMSD.cpp:219-237, NOMSD.hpp:292-331): G and
 * ov = 2 log det(PsiT Phi_w). */
int afq_mixed_density_matrix(afq_ham ham, int W, const double* d_wbuf, long ldw, double* d_G,
                             double* d_ov, void* stream);
/* NOMSD::Log_Overlap (NOMSD.cpp:280-301): ov = 2 log det(PsiT Phi_w) + log(conj c0). */
int afq_log_overlap(afq_ham ham, int W, const double* d_wbuf, long ldw, double* d_ov,
                    void* stream);
/* Real3IndexFactorization::vbias (Real3IndexFactorization.hpp:328-361):
 * v[w][n] = 2 sqrt(dt) sum_{a,k} G[w][a][k] Lnak[n][a][k]. */
int afq_vbias(afq_ham ham, int W, const double* d_G, double* d_v, void* stream);
/* Real3IndexFactorization::vHS (Real3IndexFactorization.hpp:282-326):
 * V[w][i][j] = sqrt(dt) sum_n L[i][j][n] X[w][n]. */
int afq_vhs(afq_ham ham, int W, const double* d_X, double* d_V, void* stream);
/* Real3IndexFactorization::energy (Real3IndexFactorization.hpp:138-174, 476-590) from a compact
 * G: E[w] = (E0 + 2 sum G haj, -sum_n tr(T_n^2), 2 sum_n tr(T_n)^2), T_n = G Lnak_n^T. */
int afq_energy(afq_ham ham, int W, const double* d_G, double* d_E, void* stream);
/* NOMSD::Energy (NOMSD.cpp:159-182): E and ov (= 2 log det + log conj c0) from the walkers. */
int afq_wfn_energy(afq_ham ham, int W, const double* d_wbuf, long ldw, double* d_E,
                   double* d_ov, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Propagator level (AFQMCBasePropagator.cpp, construct_X_kernel.hpp, WalkerSetUpdate.hpp).
 * ------------------------------------------------------------------------------------------- */
/* assemble_X + construct_X_impl (AFQMCBasePropagator.cpp:686-733, construct_X_kernel.hpp:33-84)
 * with N(0,1) fields d_xi [W][K] f64 in place of probit(uniform): on entry d_X holds vbias, on
 * exit X = xi + i(clamp(vbias) - clamp(vMF)); d_HW[w] = -sum vdiff (X - vdiff/2); d_mf[w] =
 * -Re(sum vMF^2)/2 + i sum X vMF (AFQMCBasePropagator.cpp:334-336). d_mf may be NULL. */
int afq_construct_X(afq_ham ham, int W, double* d_X, const double* d_xi, double* d_HW,
                    double* d_mf, void* stream);
/* Propagate with the handle's P1 (apply_propagators, AFQMCBasePropagator.cpp:216-240). */
int afq_propagate(afq_ham ham, int W, double* d_wbuf, long ldw, const double* d_V,
                  void* stream);
/* walker_update + bound_walker_weights (WalkerSetUpdate.hpp:46-218): new overlaps d_ov [W],
 * energies d_E [W][3] (read when hybrid == 0), d_mf, d_HW [W]; target population = W. */
int afq_walker_update(afq_ham ham, int W, double* d_wbuf, long ldw, const double* d_ov,
                      const double* d_E, const double* d_mf, const double* d_HW, double Eshift,
                      int hybrid, void* stream);
/* AFQMCBasePropagator::Propagate (AFQMCBasePropagator.cpp:289-392): one fused step of all W
 * walkers: G -> vbias -> X -> vHS -> P1 exp(iV) P1 -> overlap (hybrid) or energy -> weights.
 * d_xi [W][K] f64 holds the step's N(0,1) fields. Allocation-free, graph-capturable. */
int afq_step(afq_ham ham, int W, double* d_wbuf, long ldw, const double* d_xi, double Eshift,
             void* stream);
/* AFQMCBasePropagator::Orthogonalize (AFQMCBasePropagator.cpp:735-741): QR of every walker,
 * then OVLP recomputed with NOMSD::Log_Overlap. */
int afq_orthogonalize(afq_ham ham, int W, double* d_wbuf, long ldw, void* stream);
/* Back-propagation field history (WalkerSet::storeFields / advanceHistoryPos / getHistoryPos,
 * used by AFQMCBasePropagator::Propagate at AFQMCBasePropagator.cpp:341-344, 387-390): with
 * nfields > 0 every afq_step stores its assembled fields X [W][K] into slot (position % nfields) of
 * the caller's ring d_fields [W][nfields][K] c128 (the reference's Fields(nwalk, steps, nCV)) and
 * advances the position modulo the history length 3 nfields (WalkerSetBase.h:158,
 * WalkerSetBase.cpp:340); the position lives on the GPU (a captured step replays it).
 * afq_population_control moves each walker's ring row with the walker (WalkerSet::copy_walker,
 * WalkerSetBase.h:413-415). Resets the position to 0 and unregisters the weight rings;
 * nfields = 0 switches the history off. Sizes afq_back_propagate's scratch for W walkers (the
 * only allocation of back-propagation). */
int afq_set_field_history(afq_ham ham, int W, int nfields, double* d_fields, void* stream);
/* The weight rings of back-propagation (WalkerSet getWeightFactors / getWeightHistory,
 * WalkerSetBase.h:521-534), caller-owned, each [W][3 nfields] c128 for the W and nfields of the
 * registered field history; set to 1 (factors) and 0 (history) here, as resize_bp does. Every
 * afq_step then writes at the current position each walker's weight factor
 * exp(-dt (E_eff - Eshift)) / |.| / scale (0 for a walker dropped by the phaseless projection or a
 * non-finite energy) and its updated weight before the weight bound (WalkerSetUpdate.hpp:103,
 * 128-147, 184-190); afq_population_control moves both rows with the walker and writes w/|w|
 * into the last step's history entry (population_control.hpp:249-254). NULL, NULL unregisters
 * them. These are what a BackPropEstimator's path restoration reads
 * (BackPropEstimator.hpp:215-228). */
int afq_set_weight_history(afq_ham ham, double* d_factors, double* d_history, void* stream);
/* The history position (steps stored since afq_set_field_history, modulo 3 nfields).
 * Synchronises the stream. */
int afq_history_position(afq_ham ham, int* pos, void* stream);
/* AFQMCBasePropagator::BackPropagate (AFQMCBasePropagator.cpp:403-532), dense vHS: for the last
 * nbpsteps stored field sets, newest first, vHS = sqrt(dt) L X (FP64) and every reference
 * d_refs[w][nr] [M][N] (d_refs [W][nrefs][M][N] c128, the reference's Refs) propagated with the
 * adjoint propagator P1^H exp(-i vHS^H) P1^H (Propagate<'H'>, order taylor_n); after the steps
 * with ni % nstabilize == 0 and after the last one, QR (det_ops::orthogonalize) with log det R
 * added to d_logdetR [W][nrefs] c128 (zeroed first). The walkers' own Slater matrices are not
 * touched. nbpsteps <= nfields; W <= the ring's and the workspace's walker count. */
int afq_back_propagate(afq_ham ham, int W, int nbpsteps, int nstabilize, double* d_refs,
                       int nrefs, double* d_logdetR, void* stream);
/* NOMSD::Energy(wset) (NOMSD.cpp:69-80) + seedPseudoEnergy: OVLP, E1, EXX, EJ and
 * PSEUDO_ELOC = E1+EXX+EJ of every walker (execute.cpp:172-184). */
int afq_seed_properties(afq_ham ham, int W, double* d_wbuf, long ldw, void* stream);
/* population_control (population_control.hpp:219-257) on one GPU: the comb
 * (population_control.hpp:34-65) over |weight| with the given uniform in [0,1), W teeth; a
 * walker with k offspring keeps its slot and its k-1 copies fill the slots of zero-offspring
 * walkers in increasing order (swap_walkers_async, population_control.hpp:143-209); then every
 * weight becomes w/|w|. A total weight <= 1e-6 (where the reference aborts) leaves the walkers
 * untouched and is counted in afq_ham_stats. */
int afq_population_control(afq_ham ham, int W, double* d_wbuf, long ldw, double uniform,
                           void* stream);
/* Weighted sums for the estimators (EnergyEstimator.h:51-100, average_energy.hpp:25-63):
 * d_out[0] = sum w (E1+EXX+EJ), [1] = sum w, [2] = sum w exp(dt PSEUDO_ELOC), [3] = sum w
 * PSEUDO_ELOC; d_out is c128 [4]. Deterministic order. */
int afq_weighted_sums(afq_ham ham, int W, const double* d_wbuf, long ldw, double* d_out,
                      void* stream);
/* The walker sums of one EnergyEstimator::measure (EnergyEstimator.h:51-100) with
 * weight_statistics (Measurements.hpp:130-149) and the growth / mixed averages
 * (averagePseudoEnergy, averageEnergy: average_energy.hpp:53-79), in a fixed order:
 * d_out[0..2] = sum_w E_w,c w_w for c = E1, EXX, EJ (no conjugation), [3] = sum w,
 * [4] = sum exp(OVLP_w) w_w, [5] = (sum |w|, 0), [6] = (sum |w|^2, 0),
 * [7] = sum w exp(dt PSEUDO_ELOC), [8] = sum w (E1+EXX+EJ) from the walker slots; c128 [9].
 * d_E (c128 [W][3]) and d_ov (c128 [W]) replace the walkers' E1/EXX/EJ and OVLP in [0..2] and
 * [4] when given (the estimator recomputing the energy, EnergyEstimator.h:73-74); NULL takes
 * the walker slots (walkers carrying their energy). dt is the handle's time step. */
int afq_estimator_sums(afq_ham ham, int W, const double* d_wbuf, long ldw, const double* d_E,
                       const double* d_ov, double* d_out, void* stream);

/* Region timers (the AFQMCTimers regions of AFQMCTimer.h:25-63, timed on the GPU): while
 * enabled, afq_step and afq_orthogonalize record a pair of HIP events on their stream around
 * each region, up to max_records regions (0 disables). afq_ham_read_timers waits for the
 * recorded events, returns per region the summed milliseconds h_ms[AFQ_NUM_REGIONS] and the
 * number of instances h_counts[AFQ_NUM_REGIONS], and resets them. Region names:
 * afq_timer_name(i): G_for_vbias, vbias, assemble_X, vHS, propagate, overlap_G_E1, exchange,
 * walker_update, ortho. */
#define AFQ_NUM_REGIONS 9
int afq_ham_enable_timers(afq_ham ham, int max_records);
int afq_ham_read_timers(afq_ham ham, double* h_ms, long long* h_counts);
const char* afq_timer_name(int region);

/* ---------------------------------------------------------------------------------------------
 * Particle-hole multi-determinant (PHMSD) trial: energy corrections of the excited determinants
 * (detail/phmsd_impl.hpp). iexcit [ndet][2][nex] int32: the electrons (rows of refc) that are
 * excited, then the active orbitals they go to; refc [nelec] int32: the reference occupation;
 * R [nw][ndet][nex][nact] c128: the compact R matrices; wgt [ndet][nw] c128. Out-of-range
 * indices give NaN results for the walker (never an out-of-bounds access).
 * ------------------------------------------------------------------------------------------- */
/* ph_excited_2body_energy_dense_cholesky (phmsd_impl.hpp:29-147): Twina [nw][nelec][nchol][nact]
 * c128; d_EX[w] -= sum_d wgt eX_d / 2 and d_EJ[w] += sum_d wgt (KE.KE - eJ0) / 2 accumulate;
 * d_KE [ndet][nw][nchol] c128 is written. nex <= 8, nact <= 256, nelec <= 64. */
int afq_ph_excited_2body_energy(int nw, int ndet, int nex, int nact, int nelec, int nchol,
                                const int* d_iexcit, const int* d_refc, const double* d_Twina,
                                const double* d_R, const double* d_wgt, double* d_EX,
                                double* d_EJ, double* d_KE, void* stream);
/* The same computation as afq_ph_excited_2body_energy, allocating nothing. It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). afq_ph_excited_2body_energy
 * allocates a temporary for its group partials on every call (and waits for the stream before
 * freeing it); this variant keeps them in a static device buffer of
 * afq_ph_noalloc_scratch_doubles() doubles (2^20, 8 MiB) that each GPU holds until the process
 * ends, and returns AFQ_ECAPACITY when 4 * nw * ceil(nchol / 4) exceeds it. Every call on a GPU
 * shares that GPU's buffer: calls that overlap in time on one GPU (different streams, threads
 * or processes' graphs) race and must be serialised by the caller. Results equal
 * afq_ph_excited_2body_energy's bit for bit. */
int afq_ph_excited_2body_energy_noalloc(int nw, int ndet, int nex, int nact, int nelec,
                                        int nchol, const int* d_iexcit, const int* d_refc,
                                        const double* d_Twina, const double* d_R,
                                        const double* d_wgt, double* d_EX, double* d_EJ,
                                        double* d_KE, void* stream);
long afq_ph_noalloc_scratch_doubles(void);
/* ph_excited_1body_energy (phmsd_impl.hpp:149-183): S [nw][nelec][nact] c128;
 * d_E[w] += sum_d wgt sum_e (S[p_e] . R_e + S[p_e][v_e] - S[p_e][refc p_e]). */
int afq_ph_excited_1body_energy(int nw, int ndet, int nex, int nact, int nelec,
                                const int* d_iexcit, const int* d_refc, const double* d_S,
                                const double* d_R, const double* d_wgt, double* d_E, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Random numbers in step with the reference's host generators (field-for-field parity mode).
 * ------------------------------------------------------------------------------------------- */
/* split_seed (utilities/Random.cpp:7-20): splitmix64 finalizer of
 * seed << 32 | (stream & 255) << 24 | rank. The reference seeds the walker generator with
 * stream 0 and the field generator with stream 1 (execute.cpp:124-125). */
unsigned long long afq_split_seed(int seed, int rank, unsigned stream);
/* HostRandomGenerator (utilities/Random.hpp:36-49): std::mt19937 seeded with seed (kept modulo
 * 2^32, as the engine does). Host memory; destroy with afq_rng_destroy. */
int afq_rng_create(unsigned long long seed, afq_rng* out);
void afq_rng_destroy(afq_rng rng);
/* n raw 32-bit outputs of the engine into host memory. */
int afq_rng_raw(afq_rng rng, long n, unsigned* out);
/* sampleUniformFields (utilities/Random.hpp:42-48): n draws of
 * std::uniform_real_distribution<double>(0, 1) (libstdc++: two engine outputs per draw) into
 * host memory, in order. */
int afq_rng_uniform(afq_rng rng, long n, double* out);
/* probit (AFQMC/Utilities/probit.h:57-135) of n uniforms on the device: d_x[i] = probit(d_u[i]),
 * the N(0,1) fields of construct_X (construct_X_kernel.hpp:71). d_x may alias d_u. */
int afq_probit(long n, const double* d_u, double* d_x, void* stream);

/* MixedDensityMatrixForWoodbury (density_matrix.hpp:651-717), compact form: with TAB = A Phi_w
 * (A [nact][nmo] c128, the trial's active orbitals as rows) and TNN = TAB[ref rows]:
 * d_ov[w] += log det TNN (pivot signs, DBL_MIN, imaginary part wrapped to [-pi, pi]),
 * d_G [nb][nel][nmo] = TNN^-T Phi_w^T (zero if TNN is singular), d_QQ0 [nb][nact][nel] =
 * TAB TNN^-1, the T of afq_ph_overlaps / afq_ph_compact_R. Phi [nb] rows of ldw c128 holding
 * [nmo][nel]; d_work: nel * nmo c128 of device workspace. */
int afq_ph_mixed_density_matrix_woodbury(int nb, int nmo, int nel, int nact, const double* d_A,
                                         const int* d_ref, const double* d_Phi, long ldw,
                                         double* d_G, double* d_ov, double* d_QQ0,
                                         double* d_work, void* stream);
/* ph_reference_energy / ph_ref_energy_impl (Real3IndexFactorization.hpp:212-233, 598-696) for
 * one spin: with G [nw][nel][nmo] c128 (the reference determinant's compact G), Lna
 * [nchol][nact][nmo] c128 and haj [nact][nmo] c128 (NULL: no one-body term),
 * d_Twina [nw][nel][nchol][nact] = G Lnaᵀ and d_Swia [nw][nel][nact] = G hajᵀ are written, and
 * d_E [nw][3] c128 = (E0 + sum_i Swia[i][i], -scl/2 sum_{ijn} T[i][n][j] T[j][n][i],
 * 1/2 sum_n Kl[n]^2) with d_Kl [nw][nchol] = sum_i T[i][n][i]. The reference determinant
 * occupies active orbitals 0..nel-1, as the reference assumes. */
int afq_ph_reference_energy(int nw, int nel, int nmo, int nchol, int nact, double scl, double E0,
                            const double* d_G, const double* d_Lna, const double* d_haj,
                            double* d_E, double* d_Kl, double* d_Twina, double* d_Swia,
                            void* stream);
/* PHMSD trial over COLLINEAR walkers (Wavefunction.cpp:341-497 phmsd_from_params with
 * readWfn.cpp:218-340 build_ph_struct; PHMSD.hpp). Host inputs: occs [ndet][nup + ndn] int, the
 * sorted occupied orbitals of each determinant (up in [0, nmo), down in [nmo, 2 nmo)); determinant
 * 0 is the reference and must occupy orbitals 0..nup-1 and nmo..nmo+ndn-1. ci [ndet] c128.
 * orbmats [nref][nmo][nmo] c128: the trial orbitals as rows (PsiT_MO; nref 1 = one set for both
 * spins, 2 = one per spin); every orbital is active. h1 [nmo][nmo] f64, likn [nmo*nmo][nchol] f64
 * (the Cholesky vectors, as afq_ham_create takes them), E0 the constant energy. walker_block:
 * walkers per energy batch (the reference's nwalk_block_size; bounds the workspace). */
int afq_ph_create(int nmo, int nup, int ndn, int nchol, int ndet, const int* occs,
                  const double* ci, int nref, const double* orbmats, const double* h1,
                  const double* likn, double E0, int walker_block, afq_ph* out);
void afq_ph_destroy(afq_ph ph);
/* PHMSD::vMF (PHMSD.cpp:1329-1565): d_vmf [nchol] c128 = sqrt(dt) sum_s sum_{a,k} Lna_s[n][a][k]
 * G_MF,s[a][k], with G_MF the trial's mean-field density formed on the host as the reference's
 * host path forms it (configuration overlaps log det(As^H As), diagonal terms, one-orbital
 * pair terms within rows of the other spin's couplings, division by sum |c|^2 ov_up ov_dn).
 * Synchronises the stream (the host part needs no device data; the upload is ordered on it). */
int afq_ph_vmf(afq_ph ph, double dt, double* d_vmf, void* stream);
/* The PHMSD walker step (COLLINEAR walkers: rows of ldw c128 holding Phi_up [nmo][nup],
 * Phi_dn [nmo][ndn], then the AFQ_P_COUNT properties of the NOMSD layout).
 * afq_ph_prepare: the propagator for time step dt (AFQMCBasePropagator.cpp:85-157: vMF of the
 * trial with its imaginary part dropped, v0, H1 and P1 = exp(-H1/2)) and the workspace for up to
 * W walkers; the step parameters start at afq_default_params. Synchronises the stream. */
int afq_ph_prepare(afq_ph ph, double dt, int W, void* stream);
int afq_ph_set_params(afq_ph ph, const afq_step_params* p);
/* OVLP, E1/EXX/EJ and the pseudo energy of every walker from the PHMSD energy (as
 * afq_seed_properties does for NOMSD). */
int afq_ph_seed_properties(afq_ph ph, int W, double* d_wbuf, long ldw, void* stream);
/* AFQMCBasePropagator::Propagate (AFQMCBasePropagator.cpp:289-392) with the PHMSD trial:
 * PHMSD::vbias, assemble_X with the mean-field factor, vHS = sqrt(dt) L X, both spins'
 * Slater matrices propagated (P1, Taylor exp(iV), P1), the new PHMSD overlap (hybrid) or
 * PHMSD::energy, walker_update and the weight bound. d_xi [W][nchol] f64 N(0,1) fields. */
int afq_ph_step(afq_ph ph, int W, double* d_wbuf, long ldw, const double* d_xi, double Eshift,
                void* stream);
/* Orthogonalize (AFQMCBasePropagator.cpp:735-741): QR of both spins' Slater matrices, then the
 * PHMSD overlap into OVLP. */
int afq_ph_orthogonalize(afq_ph ph, int W, double* d_wbuf, long ldw, void* stream);
/* population_control (population_control.hpp:34-257) over the COLLINEAR walker layout. */
int afq_ph_population_control(afq_ph ph, int W, double* d_wbuf, long ldw, double uniform,
                              void* stream);
/* Unique excitations per spin (the reference counted as index 0) and non-empty ranks. */
int afq_ph_info(afq_ph ph, int* nunique, int* nshells);
/* PHMSD's energy algorithm (PHMSDEnergyAlgorithm, parameters.hpp:32-37; chosen at PHMSD.hpp:98
 * from the "algorithm" option, dispatched at PHMSD.cpp:89-96): AFQ_PH_ENERGY_WOODBURY (the
 * default, energy_alg1, PHMSD.cpp:377-544) or AFQ_PH_ENERGY_REFERENCE (energy_alg0,
 * PHMSD.cpp:122-370: every unique configuration's own G and energy; at most 32 electrons per
 * spin). Used by afq_ph_energy, afq_ph_seed_properties and the local-energy form of afq_ph_step. */
#define AFQ_PH_ENERGY_REFERENCE 0
#define AFQ_PH_ENERGY_WOODBURY 1
int afq_ph_set_energy_algorithm(afq_ph ph, int algorithm);
/* PHMSD::energy (PHMSD.cpp:89-96, energy_alg1 or energy_alg0): d_E [nw][3] c128 (E1, EXX, EJ of the
 * multi-determinant local energy) and d_ov [nw] c128 = log sum_D conj(c_D) <D|phi_w>. Walkers:
 * rows of ldw c128 holding Phi_up [nmo][nup] then Phi_dn [nmo][ndn] (the COLLINEAR WalkerSet). */
int afq_ph_energy(afq_ph ph, int nw, const double* d_wbuf, long ldw, double* d_E, double* d_ov,
                  void* stream);
/* PHMSD::MixedDensityMatrix (PHMSD.cpp:549-795): d_ov [nw] c128 = log sum_D conj(c_D) <D|phi_w>;
 * full = 0: d_G [nw][2 nmo][nmo] c128, the compact G of the up spin's active orbitals then the down
 * spin's (G_s = R_s^T G_ref,s / Ov); full = 1: d_G [nw][2][nmo][nmo] = A_s^T G_s. */
int afq_ph_mixed_density_matrix(afq_ph ph, int nw, const double* d_wbuf, long ldw, int full,
                                double* d_G, double* d_ov, void* stream);
/* PHMSD::vbias (PHMSD.cpp:53-72) with Real3IndexFactorization::vbias (hpp:328-361), COLLINEAR:
 * d_v [nw][nchol] c128 = sqrt(dt) sum_s sum_{a,k} Lna_s[n][a][k] G_s[w][a][k]. */
int afq_ph_vbias(afq_ph ph, int nw, const double* d_wbuf, long ldw, double dt, double* d_v,
                 void* stream);
/* The spin coupling of the determinant weights (PHMSD.cpp:441-450): C [nu][nd] (CSR: d_cp row
 * pointers [nu+1], d_ci columns, d_cv c128 CI coefficients) couples up excitation u with down
 * excitation d; d_wgt[u][w] = ratio_up[u][w] sum_d C[u][d] ratio_dn[d][w] and
 * d_ov[w] = sum_u d_wgt[u][w] (the multi-determinant overlap over the reference's).
 * ratios [n][nw] c128 as afq_ph_overlaps writes them (row 0 = the reference, ratio 1). */
int afq_ph_spin_weights(int nw, int nu, int nd, const int* d_cp, const int* d_ci,
                        const double* d_cv, const double* d_rup, const double* d_rdn,
                        double* d_wgt, double* d_ov, void* stream);
/* calculate_overlaps (phmsd_helpers.hpp:111-195) for the rank-nex excitations of one spin:
 * d_ov[d][w] = det A, A(p, q) = T[w][e_(nex+p)][e_q] with e = iexcit[d]; T [nw][nact][nelec]
 * c128 (the walker's active-space matrix); 0 for a singular A. nex <= 8. */
int afq_ph_overlaps(int nw, int ndet, int nex, int nact, int nelec, const int* d_iexcit,
                    const double* d_T, double* d_ov, void* stream);
/* get_compact_ph_R_matrices (phmsd_helpers.hpp:317-383): d_R [nw][ndet][nex][nact] c128,
 * R(p, orbs(i)) = sum_q Q(p, q) (delta(i, e_q) - T[e_(nex+q)][i]) with Q = A^-1 and orbs the
 * excited configuration (refc with e_q -> e_(nex+q)); zero elsewhere and where A is singular. */
int afq_ph_compact_R(int nw, int ndet, int nex, int nact, int nelec, const int* d_iexcit,
                     const int* d_refc, const double* d_T, double* d_R, void* stream);
/* calculate_R (phmsd_helpers.hpp:200-262) for one shell, accumulated into d_R [nw][nelec][nact]
 * c128: R[w][i][orbs_d(i)] += wgt[d][w] and R[w][e_p][:] += wgt[d][w] Rc[w][d][p][:] for every
 * determinant with d_ov[d][w] != 0, in determinant order; nex = 0 adds the reference
 * determinant's R[w][i][refc(i)] += wgt[0][w]. Call it for nex = 0, 1, 2, ... in turn. */
int afq_ph_accumulate_R(int nw, int ndet, int nex, int nact, int nelec, const int* d_iexcit,
                        const int* d_refc, const double* d_Rc, const double* d_ov,
                        const double* d_wgt, double* d_R, void* stream);

/* Lane layout of the MFMA builtins the kernels use, for a self-check; d_out [64][4] f64
 * receives accumulator element r of every lane.
 * kind 0 / 1: FP64 / FP32 16x16x4, D = A B with A[i][k] = i + 100 k, B[k][j] = (k == j % 4).
 * kind 2 / 3: BF16 16x16x32, D = A B with A[i][k] = i + 1 and B[k][j] = j + 1 only for k = 0
 *             (kind 2) or k = 31 (kind 3), so D[i][j] = (i + 1)(j + 1).
 * kind 4 / 5: int8 16x16x64 (i32 accumulator), as kinds 2 / 3 with k = 0 or k = 63. */
int afq_probe_mfma(int kind, double* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SAFIRE_AMD_AFQ_H */
