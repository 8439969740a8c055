// Internal declarations of the MI355X walker-step library: device complex helpers, the MFMA
// wrappers, the per-system engine state and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/afq.h"

namespace afq {

// ------------------------------------------------------------------ error plumbing
void set_error(const std::string& msg);
struct Error {
  int code;
  std::string msg;
};

#define AFQ_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::afq::Error{AFQ_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)};      \
  } while (0)

#define AFQ_CHECK(cond, code, msg)                                                          \
  do {                                                                                      \
    if (!(cond)) throw ::afq::Error{(code), (msg)};                                         \
  } while (0)

// ------------------------------------------------------------------ complex on the device
struct cd {
  double re, im;
};

__host__ __device__ inline cd cmk(double r, double i) { return cd{r, i}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cd operator-(cd a) { return cd{-a.re, -a.im}; }
__host__ __device__ inline cd operator*(cd a, cd b) {
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__host__ __device__ inline cd& operator+=(cd& a, cd b) {
  a.re += b.re;
  a.im += b.im;
  return a;
}
__host__ __device__ inline cd& operator-=(cd& a, cd b) {
  a.re -= b.re;
  a.im -= b.im;
  return a;
}
__host__ __device__ inline cd cconj(cd a) { return cd{a.re, -a.im}; }
__host__ __device__ inline double cnorm(cd a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ inline double cabs_(cd a) { return hypot(a.re, a.im); }
__host__ __device__ inline double cabs1(cd a) { return fabs(a.re) + fabs(a.im); }
__host__ __device__ inline bool cis0(cd a) { return a.re == 0.0 && a.im == 0.0; }
// complex division with Smith's scaling (the -fcx-fortran-rules division of the host build)
__host__ __device__ inline cd cdiv(cd a, cd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cd{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.im + b.re * r;
    return cd{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
__host__ __device__ inline cd clog_(cd a) { return cd{log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }
__host__ __device__ inline cd cexp_(cd a) {
  double e = exp(a.re);
  return cd{e * cos(a.im), e * sin(a.im)};
}

// ------------------------------------------------------------------ MFMA wrappers (gfx950)
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16). Lane l supplies A[l%16][l/16] and B[l/16][l%16]; accumulator
// element r of lane l is D[acc_row<R>(l/16, r)][l%16]. The row maps were measured with the
// probe kernel (afq_probe_mfma): FP64 holds rows g, g+4, g+8, g+12; FP32 holds rows 4g..4g+3.
__device__ inline d4 mfma16x16x4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma16x16x4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// D(16x16, f32) += A(16x32, bf16) B(32x16, bf16): lane l supplies A[l%16][8(l/16)+q] and
// B[8(l/16)+q][l%16], q = 0..7; the accumulator follows the FP32 layout (acc_row<float>).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ inline f4 mfma_bf16(bf16x8 a, bf16x8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <typename R> __host__ __device__ inline int acc_row(int g, int r);
template <> __host__ __device__ inline int acc_row<double>(int g, int r) { return 4 * r + g; }
template <> __host__ __device__ inline int acc_row<float>(int g, int r) { return 4 * g + r; }

template <typename R> struct acc_of;
template <> struct acc_of<double> { typedef d4 type; };
template <> struct acc_of<float> { typedef f4 type; };

// ------------------------------------------------------------------ engine state
struct Workspace {
  int cap = 0;
  double* G = nullptr;       // c128 [W][N][M]
  double* X = nullptr;       // c128 [W][K]
  double* V = nullptr;       // c128 [W][M][M]
  double* HW = nullptr;      // c128 [W]
  double* mf = nullptr;      // c128 [W]
  double* ov = nullptr;      // c128 [W]
  double* E = nullptr;       // c128 [W][3]
  double* Kl = nullptr;      // c128 [W][K]
  double* exx = nullptr;     // c128 [W][nchunk]
  double* cum = nullptr;     // f64 [W]   population control: cumulative |w|
  int* off = nullptr;        // i32 [W]   offspring counts
  int* src = nullptr;        // i32 [W]   source walker of each slot (-1 = keep)
  int* ok = nullptr;         // i32 [1]   comb accepted the population
  double* split = nullptr;   // c128 [kMaxSplits][W][K]   split-k partials of vbias
  long split_elems = 0;
};

// Region timers of a step (the AFQMCTimers regions of AFQMCTimer.h:25-63), as HIP events
// recorded on the launch stream around each region; read back (and summed) on demand.
enum Region {
  R_G_FOR_VBIAS = 0, R_VBIAS, R_ASSEMBLE_X, R_VHS, R_PROPAGATE, R_OVERLAP_G_E1, R_EXCHANGE,
  R_WALKER_UPDATE, R_ORTHO, R_COUNT
};
struct Timers {
  bool on = false;
  int cap = 0, used = 0;
  std::vector<hipEvent_t> ev;      // 2 per record
  std::vector<int> region;         // region of record i
  double ms[R_COUNT] = {0};
  long long count[R_COUNT] = {0};
};
// RAII region scope: records start/stop events when the timers are on.
struct RegionScope {
  Timers& t;
  hipStream_t s;
  int rec = -1;
  RegionScope(Timers& t_, hipStream_t s_, int r);
  ~RegionScope();
};

}  // namespace afq

struct afq_ham_s {
  int M = 0, N = 0, K = 0, walker_type = AFQ_CLOSED;
  double dt = 0.005, E0 = 0.0;
  double ci0_re = 1.0, ci0_im = 0.0;
  int precision = AFQ_FP64;
  int cfg_vbias = 0, cfg_vhs = 0;   // GEMM tilings (afq_autotune)
  afq_step_params params;
  int device = 0;
  // device tensors (built on the GPU at create)
  double* PsiT = nullptr;    // c128 [N][M]
  double* haj = nullptr;     // c128 [N][M]
  double* Lnak = nullptr;    // c128 [K][N][M]
  double* Likn = nullptr;    // f64 [M*M][K]
  double* h = nullptr;       // f64 [M][M]
  double* v0 = nullptr;      // f64 [M][M]
  double* vMF = nullptr;     // c128 [K]
  double* P1 = nullptr;      // c128 [M][M]
  double* H1 = nullptr;      // c128 [M][M]
  float* Lnak_f = nullptr;   // c64 [K][N][M]   (FP32 contraction copies)
  float* Likn_f = nullptr;   // f32 [M*M][K]
  __bf16* Lnak_b = nullptr;  // bf16 [6 planes][K][16 NT][32 KC]  (BF16X3 exchange)
  __bf16* Lnak_g = nullptr;  // bf16 [6 planes][K][N*M padded to 32]   (BF16X3 vbias)
  __bf16* Likn_g = nullptr;  // bf16 [2 planes][M*M][K padded to 32]   (BF16X3 vHS)
  int8_t* Lnak_q = nullptr;  // i8 [5 or 6 slices][re,im][K][N][Mp] (INT8_SPLIT exchange)
  double* Lnak_u = nullptr;  // f64 [K][N]   power-of-two unit of each Lnak row
  int* Lnak_ord = nullptr;   // i32 [K]      slice-product order of each vector (N > 16)
  // [7]: eloc total/upper/lower, weight total/upper, population control calls/collapses
  unsigned long long* stats = nullptr;
  // back-propagation: the caller's field ring [fields_W][nfields][K] c128 and its position
  double* fields = nullptr;
  int nfields = 0, fields_W = 0;
  // weight factors and weight history [fields_W][3 nfields] c128 (caller-owned; null = none)
  double* wfac = nullptr;
  double* whis = nullptr;
  int* hist_pos = nullptr;   // device int
  double* P1H = nullptr;     // c128 [M][M]
  double* bp_ldet = nullptr; // c128 [bp_cap]
  int bp_cap = 0;
  afq::Workspace ws;
  afq::Timers timers;
};

namespace afq {

// Raises a kernel's dynamic-LDS limit to `bytes` (when above 64 KB) once per device. The
// attribute call is not a stream operation; doing it once keeps it out of every later launch,
// including launches recorded into a graph.
void ensure_dynamic_lds(const void* kern, size_t bytes);

// ------------------------------------------------------------------ launchers
// k_green.hip
void launch_green(hipStream_t s, int nb, int M, int N, const double* PsiT, const double* Phi,
                  long ldw, double* G, double* ov, int ov_mode, double ov_scale, cd ov_add,
                  const double* haj, double E0, double e1_scale, double* E, long ldE);
// k_gemm.hip:  C[m][n] = alpha * sum_k A[m][k] B[n][k]; A complex, B complex or real.
// split_ws (c128, split_ws_elems complex numbers) enables a deterministic split of k; NULL = none.
constexpr int kMaxSplits = 8;
int gemm_splits(int m, int n, int k);
// tilings launch_gemm_nt_z can run for a precision and kind of B (the cfg argument: 0 .. n-1)
int gemm_tile_configs(int precision, bool b_complex);
void launch_gemm_nt_z(hipStream_t s, int precision, int m, int n, int k, double alpha,
                      const double* A, long lda, const void* B, long ldb, bool b_complex,
                      double* C, long ldc, double* split_ws = nullptr, long split_ws_elems = 0,
                      int cfg = 0);
// k_energy.hip
int device_cus();                        // compute units of the current device (cached)
int exx_nchunks(int W, int N, int K);   // chunks of Cholesky vectors the exchange kernel uses
void launch_split_reduce(hipStream_t s, int m, int n, int splits, double alpha,
                         const double* split_ws, double* C, long ldc);
// k_gemm_bf16.hip: the BF16X3 form of the Cholesky GEMMs (three-product complex multiplication,
// hi/lo BF16 splits, FP32 accumulation); B as BF16 planes from split_planes_bf16 (complex: 6
// planes, real: 2, rows padded to bf16_row(cols))
int bf16_row(int cols);
void split_planes_bf16(hipStream_t s, long rows, int cols, bool complex_src, const double* src,
                       __bf16* dst);
void launch_gemm_bf16(hipStream_t s, int m, int n, int k, double alpha, const double* A, long lda,
                      const __bf16* Bp, bool b_complex, double* C, long ldc,
                      double* split_ws = nullptr, long split_ws_elems = 0);
// k_exchange_bf16.hip: the exchange in the BF16X3 form (N <= 32, M <= 128)
bool exchange_bf16_supported(int M, int N);
size_t exchange_bf16_elems(int M, int N, int K);   // bf16 elements of the exchange planes
int exchange_bf16_chunk(int W, int N, int K);
int exchange_bf16_chunks(int W, int N, int K);
void split_lnak_bf16(hipStream_t s, int M, int N, int K, const double* Lnak, __bf16* Lb);
void launch_exchange_bf16(hipStream_t s, int W, int M, int N, int K, const double* G,
                          const __bf16* Lb, double* Kl, double* exx_part);
// k_exchange_i8.hip: the exchange on int8 MFMA, Ozaki-split rows (N <= 16, M <= 128)
bool exchange_i8_supported(int M, int N);
int exchange_i8_row(int M);                        // Mp: padded row length of the slices
size_t exchange_i8_bytes(int M, int N, int K);     // size of the Lnak slices
int exchange_i8_chunks(int W, int N, int K);   // tr(T^2) partials per walker
void split_lnak_i8(hipStream_t s, int M, int N, int K, const double* Lnak, int8_t* Lq,
                   double* Lu);
int exchange_i8_slices(int N);                     // slices of each Lnak row
// per Cholesky vector the slice-product order of the 17..32-electron kernel (set-up, synchronous)
void exchange_i8_orders(hipStream_t s, int M, int N, int K, const double* Lnak, int* ord);
void launch_exchange_i8(hipStream_t s, int W, int M, int N, int K, const double* G,
                        const int8_t* Lq, const double* Lu, const int* ord, double* Kl,
                        double* exx_part);
void launch_exchange(hipStream_t s, int precision, int W, int M, int N, int K, const double* G,
                     const void* Lnak, double* Kl, double* exx_part);
void launch_energy_finalize(hipStream_t s, int W, int K, int nchunk, int walker_type,
                            const double* Kl, const double* exx_part, double* E);
void launch_e1(hipStream_t s, int W, int M, int N, const double* G, const double* haj,
               double E0, double scl, double* E);
// k_propagate.hip
void launch_propagate(hipStream_t s, int nb, int M, int N, double* Phi, long ldw,
                      const double* P1, const double* V, int order);
// k_qr.hip
void launch_orthogonalize(hipStream_t s, int nb, int M, int N, double* Phi, long ldw,
                          double* logdetR);
// k_walker.hip
void launch_store_fields(hipStream_t s, int W, int K, const double* X, double* ring, int nfields,
                         const int* pos);
void launch_advance_history(hipStream_t s, int* pos, int nhist);
void launch_neg_adjoint(hipStream_t s, int W, int M, double* V);
void launch_adjoint(hipStream_t s, int M, const double* A, double* B);
void launch_add_strided(hipStream_t s, int W, const double* x, double* out, long stride);
void launch_construct_X(hipStream_t s, int W, int K, double vbias_bound, int project,
                        int free_projection, const double* vMF, const double* xi, double* X,
                        double* HW, double* mf);
void launch_walker_update(hipStream_t s, int W, double* wbuf, long ldw, long poff,
                          const afq_step_params& p, double dt, double Eshift, double E0,
                          const double* ov, const double* E, const double* mf,
                          const double* HW, int hybrid, unsigned long long* stats,
                          double* wfac = nullptr, double* whis = nullptr, int nhist = 0,
                          const int* pos = nullptr);
void launch_set_properties(hipStream_t s, int W, double* wbuf, long ldw, long poff,
                           const double* ov, const double* E, int seed_pseudo);
void launch_weighted_sums(hipStream_t s, int W, const double* wbuf, long ldw, long poff,
                          double dt, double* out);
void launch_estimator_sums(hipStream_t s, int W, const double* wbuf, long ldw, long poff,
                           double dt, const double* E, const double* ov, double* out);
void launch_probe_mfma(hipStream_t s, int kind, double* out);
// the per-walker back-propagation rows that population control moves with the walkers: the
// field ring [W][fields_ld] c128 and the weight-factor / weight-history rings [W][nhist] c128
// (null when absent), and the history position on the device
struct BackPropRows {
  double* fields = nullptr;
  long fields_ld = 0;
  double* wfac = nullptr;
  double* whis = nullptr;
  int nhist = 0;
  const int* pos = nullptr;
};
void launch_population_control(hipStream_t s, int W, double* wbuf, long ldw, long poff,
                               double r, double* cum, int* off, int* src, int* ok,
                               unsigned long long* stats, const BackPropRows* bp = nullptr);
// k_setup.hip
void setup_hamiltonian(afq_ham_s* H, hipStream_t s);
void launch_v0(hipStream_t s, int M, int K, const double* Likn, double* v0);
void launch_zero_imag(hipStream_t s, int n, double* v);
// P1 [M][M] c128 on the host as the reference applies it: |x| <= 1e-8 dropped, imaginary
// roundoff dropped when P1 is real to 1e-14
void sparsify_one_body_propagator(int M, double* P);
void build_one_body_propagator(hipStream_t s, int M, int K, double dt, const double* h,
                               const double* v0, const double* Likn, const double* vMF,
                               double* H1, double* P1);
// k_phmsd.hip
// ws: ph_2body_workspace() doubles for the staged kernel (NULL or too small: the per-walker one)
long ph_2body_workspace(int nw, int nact, int nelec, int nchol);
void launch_ph_2body(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec, int nchol,
                     const int* iexcit, const int* refc, const double* Twina, const double* R,
                     const double* wgt, double* EX, double* EJ, double* KE,
                     double* ws = nullptr, long ws_doubles = 0);
void launch_ph_1body(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                     const int* iexcit, const int* refc, const double* S, const double* R,
                     const double* wgt, double* E);
void launch_ph_woodbury(hipStream_t s, int nb, int nmo, int nel, int nact, const double* A,
                        const int* ref, const double* Phi, long ldw, double* G, double* ov,
                        double* QQ0, double* work);
void launch_ph_spin_weights(hipStream_t s, int nw, int nu, const int* cp, const int* ci,
                            const double* cv, const double* rup, const double* rdn, double* wgt,
                            double* ov);
void launch_ph_overlaps(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                        const int* iexcit, const double* T, double* ov);
void launch_ph_compact_R(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                         const int* iexcit, const int* refc, const double* T, double* R);
void launch_ph_reference_energy(hipStream_t s, int nw, int nel, int nmo, int nchol, int nact,
                                double scl, double E0, const double* G, const double* Lna,
                                const double* haj, double* E, double* Kl, double* Twina,
                                double* Swia);
void launch_ph_couple_ke(hipStream_t s, int nw, int nchol, int nout, const int* cp, const int* ci,
                         const double* cv, const double* rin, const double* KEin,
                         const double* rout, double* KEout);
void launch_ph_ke_dot(hipStream_t s, int nw, int nchol, int nd, const double* a, const double* b,
                      double* out, int ostride);
void launch_ph_accumulate(hipStream_t s, int nw, const double* Eref, const double* Ov,
                          const double* Ex, double* acc);
void launch_ph_finalize(hipStream_t s, int nw, const double* acc, const double* Ov,
                        const double* log_ov, double* E, double* ov_out);
void launch_ph_fill(hipStream_t s, long n, double re, double im, double* x);
void launch_ph_rg(hipStream_t s, int nw, int nel, int nact, int nmo, int ldrows, int row0,
                  const double* R, const double* GA, const double* Ov, double* G);
void launch_ph_full_g(hipStream_t s, int nw, int nact, int nmo, const double* A0,
                      const double* A1, const double* Gc, double* Gf);
void launch_ph_log_overlap(hipStream_t s, int nw, const double* Ov, const double* log_ov,
                           double* out);
void launch_ph0_products(hipStream_t s, int nw, int nmo, int nel, int nact, int nchol,
                         const double* Phi, long ldw, const double* A, const double* haj,
                         const double* Lna, double* Phit, double* TAB, double* Hp, double* P);
void launch_ph0_config(hipStream_t s, int nw, int nc, int nel, int nact, int nchol,
                       const int* conf, const double* TAB, const double* Hp, const double* P,
                       double E0, double* ld, double* Emsd, double* KE, const double* KEr,
                       double* Tinv);
void launch_ph0_ratios(hipStream_t s, int nw, int nc, double* r, double* log_ov);
void launch_ph0_assemble(hipStream_t s, int nw, int nu, int nd, const double* wgt0,
                         const double* Em0, const double* wgt1, const double* Em1,
                         const double* r1, double* acc);
void launch_ph_accumulate_R(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                            const int* iexcit, const int* refc, const double* Rc,
                            const double* ov, const double* wgt, double* R);

}  // namespace afq
