// C ABI of the walker-step library (declared in include/afq.h). Host C++: argument checks,
// device memory of a system, and the launch sequence of one step. Every entry point converts
// failures into an afq_status; the message is kept per thread for afq_last_error().
#include "afq_internal.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cctype>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace afq {

void ensure_dynamic_lds(const void* kern, size_t bytes) {
  if (bytes <= 64 * 1024) return;
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> done;   // (device, kernel) -> bytes
  int dev = 0;
  AFQ_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  size_t& cur = done[{dev, kern}];
  if (cur >= bytes) return;
  AFQ_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  cur = bytes;
}


namespace {
thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
  try {
    f();
    return AFQ_OK;
  } catch (const Error& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return AFQ_ENOMEM;
  } catch (...) {
    g_err = "unknown error";
    return AFQ_EINVAL;
  }
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

template <typename T>
void dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e != hipSuccess)
    throw Error{AFQ_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)};
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_workspace(Workspace& ws) {
  dfree(ws.G);
  dfree(ws.X);
  dfree(ws.V);
  dfree(ws.HW);
  dfree(ws.mf);
  dfree(ws.ov);
  dfree(ws.E);
  dfree(ws.Kl);
  dfree(ws.exx);
  dfree(ws.split);
  ws.split_elems = 0;
  dfree(ws.cum);
  dfree(ws.off);
  dfree(ws.src);
  dfree(ws.ok);
  ws.cap = 0;
}

void check_ham(afq_ham H) { AFQ_CHECK(H != nullptr, AFQ_EINVAL, "null afq_ham"); }

void check_walkers(afq_ham H, int W, long ldw) {
  AFQ_CHECK(W >= 0, AFQ_EBADSHAPE, "negative walker count");
  AFQ_CHECK(W <= H->ws.cap, AFQ_ECAPACITY,
            "more walkers than afq_ham_reserve() sized the workspace for");
  AFQ_CHECK(ldw >= (long)H->M * H->N + AFQ_P_COUNT, AFQ_EBADSHAPE,
            "walker stride smaller than M*N + 11");
}

cd ci_log(afq_ham H) {
  const double a = std::hypot(H->ci0_re, H->ci0_im);
  if (a < 1e-12) return cmk(std::log(1e-12), 0.0);
  return clog_(cmk(H->ci0_re, -H->ci0_im));
}

// vbias / vHS operands: FP64 in the plain form, the FP32 copies in both reduced forms
const void* L_for(afq_ham H) { return H->precision == AFQ_FP64 ? (const void*)H->Lnak : (const void*)H->Lnak_f; }
const void* Likn_for(afq_ham H) { return H->precision == AFQ_FP64 ? (const void*)H->Likn : (const void*)H->Likn_f; }
double scl(afq_ham H) { return H->walker_type == AFQ_CLOSED ? 2.0 : 1.0; }

// vbias = 2 sqrt(dt) G Lnak^T and vHS = sqrt(dt) L X in the handle's precision (the BF16X3 form
// on its own GEMM, the others on the FP64/FP32 MFMA GEMM with tiling cfg)
void run_vbias(afq_ham H, hipStream_t s, int W, const double* G, double* v, int cfg) {
  const long NM = (long)H->N * H->M;
  const double alpha = std::sqrt(H->dt) * scl(H);
  if (H->precision == AFQ_BF16X3)
    launch_gemm_bf16(s, W, H->K, (int)NM, alpha, G, NM, H->Lnak_g, true, v, H->K, H->ws.split,
                     H->ws.split_elems);
  else
    launch_gemm_nt_z(s, H->precision, W, H->K, (int)NM, alpha, G, NM, L_for(H), NM, true, v,
                     H->K, H->ws.split, H->ws.split_elems, cfg);
}

void run_vhs(afq_ham H, hipStream_t s, int W, const double* X, double* V, int cfg) {
  const int MM = H->M * H->M;
  if (H->precision == AFQ_BF16X3)
    launch_gemm_bf16(s, W, MM, H->K, std::sqrt(H->dt), X, H->K, H->Likn_g, false, V, MM);
  else
    launch_gemm_nt_z(s, H->precision, W, MM, H->K, std::sqrt(H->dt), X, H->K, Likn_for(H),
                     H->K, false, V, MM, nullptr, 0, cfg);
}

// the exchange contraction in the handle's precision: Kl and per-chunk tr(T^2) partials
void run_exchange(afq_ham H, hipStream_t s, int W, const double* G) {
  if (H->precision == AFQ_BF16X3)
    launch_exchange_bf16(s, W, H->M, H->N, H->K, G, H->Lnak_b, H->ws.Kl, H->ws.exx);
  else if (H->precision == AFQ_INT8_SPLIT)
    launch_exchange_i8(s, W, H->M, H->N, H->K, G, H->Lnak_q, H->Lnak_u, H->Lnak_ord, H->ws.Kl,
                       H->ws.exx);
  else
    launch_exchange(s, H->precision, W, H->M, H->N, H->K, G, L_for(H), H->ws.Kl, H->ws.exx);
}

int exchange_chunks(afq_ham H, int W) {
  if (H->precision == AFQ_BF16X3) return exchange_bf16_chunks(W, H->N, H->K);
  if (H->precision == AFQ_INT8_SPLIT) return exchange_i8_chunks(W, H->N, H->K);
  return exx_nchunks(W, H->N, H->K);
}

void finalize_energy(afq_ham H, hipStream_t s, int W, double* E) {
  launch_energy_finalize(s, W, H->K, exchange_chunks(H, W), H->walker_type, H->ws.Kl,
                         H->ws.exx, E);
}

// G, ov (= 2 logdet + log conj c0) and E (E1 by the Green's function kernel, EXX/EJ by the
// exchange kernel) of all walkers into the workspace (or E/ov given)
void energy_from_walkers(afq_ham H, hipStream_t s, int W, const double* wbuf, long ldw,
                         double* E, double* ov) {
  const int M = H->M, N = H->N;
  launch_green(s, W, M, N, H->PsiT, wbuf, ldw, H->ws.G, ov, 1, scl(H), ci_log(H), H->haj,
               H->E0, scl(H), E, 3);
  run_exchange(H, s, W, H->ws.G);
  finalize_energy(H, s, W, E);
}

const char* const kRegionNames[R_COUNT] = {"G_for_vbias", "vbias",          "assemble_X",
                                           "vHS",         "propagate",      "overlap_G_E1",
                                           "exchange",    "walker_update",  "ortho"};

void destroy_timers(Timers& t) {
  for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  t.ev.clear();
  t.region.clear();
  t.cap = t.used = 0;
  t.on = false;
}

}  // namespace

RegionScope::RegionScope(Timers& t_, hipStream_t s_, int r) : t(t_), s(s_) {
  if (t.on && t.used < t.cap) {
    rec = t.used++;
    t.region[rec] = r;
    (void)hipEventRecord(t.ev[2 * rec], s);
  }
}
RegionScope::~RegionScope() {
  if (rec >= 0) (void)hipEventRecord(t.ev[2 * rec + 1], s);
}

void set_error(const std::string& msg) { g_err = msg; }

}  // namespace afq

using namespace afq;

extern "C" {

void afq_default_params(afq_step_params* p) {
  if (!p) return;
  p->taylor_n = 6;
  p->vbias_bound = 50.0;
  p->upper_cutoff_scale = 10.0;
  p->lower_cutoff_scale = 1.0;
  p->weight_bound_floor = 100.0;
  p->weight_bound_fraction = 0.1;
  p->hybrid = 1;
  p->free_projection = 0;
  p->use_cp_constraint = 0;
  p->project_force_bias = 0;
}

const char* afq_last_error(void) { return g_err.c_str(); }
int afq_version(void) { return 1; }

int afq_ham_create(int M, int N, int K, const double* h_h, const double* h_Likn,
                   const double* h_PsiT, double E0, double dt, int walker_type, afq_ham* out) {
  return guard([&] {
    AFQ_CHECK(out != nullptr, AFQ_EINVAL, "null output handle");
    *out = nullptr;
    AFQ_CHECK(M > 0 && N > 0 && K > 0 && N <= M, AFQ_EBADSHAPE, "bad (M, N, K)");
    AFQ_CHECK(h_h && h_Likn && h_PsiT, AFQ_EINVAL, "null input array");
    AFQ_CHECK(walker_type == AFQ_CLOSED, AFQ_EUNSUPPORTED,
              "only CLOSED walkers are supported by afq_ham_create");
    AFQ_CHECK(N <= 32, AFQ_EUNSUPPORTED, "more than 32 electrons per spin");
    AFQ_CHECK(dt > 0, AFQ_EINVAL, "dt must be positive");
    afq_ham H = new afq_ham_s();
    try {
      H->M = M;
      H->N = N;
      H->K = K;
      H->dt = dt;
      H->E0 = E0;
      H->walker_type = walker_type;
      afq_default_params(&H->params);
      AFQ_HIP(hipGetDevice(&H->device));
      const size_t MM = (size_t)M * M, NM = (size_t)N * M;
      dalloc(&H->PsiT, 2 * NM);
      dalloc(&H->haj, 2 * NM);
      dalloc(&H->Lnak, 2 * NM * K);
      dalloc(&H->Likn, MM * K);
      dalloc(&H->h, MM);
      dalloc(&H->v0, MM);
      dalloc(&H->vMF, 2 * (size_t)K);
      dalloc(&H->P1, 2 * MM);
      dalloc(&H->H1, 2 * MM);
      dalloc(&H->Lnak_f, 2 * NM * K);
      dalloc(&H->Likn_f, MM * K);
      dalloc(&H->stats, 7);
      AFQ_HIP(hipMemcpy(H->PsiT, h_PsiT, 2 * NM * sizeof(double), hipMemcpyHostToDevice));
      AFQ_HIP(hipMemcpy(H->Likn, h_Likn, MM * K * sizeof(double), hipMemcpyHostToDevice));
      AFQ_HIP(hipMemcpy(H->h, h_h, MM * sizeof(double), hipMemcpyHostToDevice));
      AFQ_HIP(hipMemset(H->stats, 0, 7 * sizeof(unsigned long long)));
      setup_hamiltonian(H, nullptr);
      if (exchange_bf16_supported(M, N)) {
        dalloc(&H->Lnak_b, exchange_bf16_elems(M, N, K));
        split_lnak_bf16(nullptr, M, N, K, H->Lnak, H->Lnak_b);
        dalloc(&H->Lnak_g, 6 * (size_t)K * bf16_row((int)NM));
        split_planes_bf16(nullptr, K, (int)NM, true, H->Lnak, H->Lnak_g);
        dalloc(&H->Likn_g, 2 * MM * bf16_row(K));
        split_planes_bf16(nullptr, (long)MM, K, false, H->Likn, H->Likn_g);
      }
      if (exchange_i8_supported(M, N)) {
        dalloc(&H->Lnak_q, exchange_i8_bytes(M, N, K));
        dalloc(&H->Lnak_u, (size_t)K * N);
        split_lnak_i8(nullptr, M, N, K, H->Lnak, H->Lnak_q, H->Lnak_u);
        if (N > 16) {
          dalloc(&H->Lnak_ord, (size_t)K);
          exchange_i8_orders(nullptr, M, N, K, H->Lnak, H->Lnak_ord);
        }
      }
      AFQ_HIP(hipDeviceSynchronize());
    } catch (...) {
      afq_ham_destroy(H);
      throw;
    }
    *out = H;
  });
}

int afq_ham_destroy(afq_ham H) {
  return guard([&] {
    if (!H) return;
    destroy_timers(H->timers);
    free_workspace(H->ws);
    dfree(H->PsiT);
    dfree(H->haj);
    dfree(H->Lnak);
    dfree(H->Likn);
    dfree(H->h);
    dfree(H->v0);
    dfree(H->vMF);
    dfree(H->P1);
    dfree(H->H1);
    dfree(H->Lnak_f);
    dfree(H->Likn_f);
    dfree(H->Lnak_ord);
    dfree(H->Lnak_b);
    dfree(H->Lnak_g);
    dfree(H->Likn_g);
    dfree(H->Lnak_q);
    dfree(H->Lnak_u);
    dfree(H->stats);
    dfree(H->hist_pos);
    dfree(H->P1H);
    dfree(H->bp_ldet);
    delete H;
  });
}

int afq_ham_reserve(afq_ham H, int max_walkers) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(max_walkers >= 0, AFQ_EBADSHAPE, "negative walker count");
    if (max_walkers <= H->ws.cap) return;
    AFQ_HIP(hipDeviceSynchronize());
    free_workspace(H->ws);
    const size_t W = max_walkers;
    const size_t M = H->M, N = H->N, K = H->K;
    Workspace& ws = H->ws;
    dalloc(&ws.G, 2 * W * N * M);
    dalloc(&ws.X, 2 * W * K);
    dalloc(&ws.V, 2 * W * M * M);
    dalloc(&ws.HW, 2 * W);
    dalloc(&ws.mf, 2 * W);
    dalloc(&ws.ov, 2 * W);
    dalloc(&ws.E, 6 * W);
    dalloc(&ws.Kl, 2 * W * K);
    dalloc(&ws.exx, 2 * W * (size_t)(K + 2));   // at most K + 1 tr(T^2) partials per walker
    ws.split_elems = (long)kMaxSplits * W * K;
    dalloc(&ws.split, 2 * (size_t)ws.split_elems);
    dalloc(&ws.cum, W);
    dalloc(&ws.off, W);
    dalloc(&ws.src, W);
    dalloc(&ws.ok, 1);
    ws.cap = max_walkers;
  });
}

int afq_ham_set_precision(afq_ham H, int precision) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(precision >= AFQ_FP64 && precision <= AFQ_INT8_SPLIT, AFQ_EINVAL,
              "unknown precision");
    AFQ_CHECK(precision != AFQ_BF16X3 || H->Lnak_b != nullptr, AFQ_EUNSUPPORTED,
              "BF16X3 form needs N <= 32 and M <= 128");
    AFQ_CHECK(precision != AFQ_INT8_SPLIT || H->Lnak_q != nullptr, AFQ_EUNSUPPORTED,
              "INT8_SPLIT exchange needs N <= 32 and M <= 128");
    H->precision = precision;
    H->cfg_vbias = H->cfg_vhs = 0;   // tilings are tuned per precision (afq_autotune)
  });
}

int afq_ham_set_params(afq_ham H, const afq_step_params* p) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(p != nullptr, AFQ_EINVAL, "null params");
    AFQ_CHECK(p->taylor_n >= 1, AFQ_EINVAL, "taylor_n must be >= 1");
    H->params = *p;
  });
}

int afq_ham_get_params(afq_ham H, afq_step_params* p) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(p != nullptr, AFQ_EINVAL, "null params");
    *p = H->params;
  });
}

int afq_ham_tensors(afq_ham H, double* haj, double* Lnak, double* v0, double* vMF, double* P1,
                    double* H1) {
  return guard([&] {
    check_ham(H);
    const size_t MM = (size_t)H->M * H->M, NM = (size_t)H->N * H->M;
    AFQ_HIP(hipDeviceSynchronize());
    if (haj) AFQ_HIP(hipMemcpy(haj, H->haj, 2 * NM * sizeof(double), hipMemcpyDeviceToHost));
    if (Lnak)
      AFQ_HIP(hipMemcpy(Lnak, H->Lnak, 2 * NM * H->K * sizeof(double), hipMemcpyDeviceToHost));
    if (v0) AFQ_HIP(hipMemcpy(v0, H->v0, MM * sizeof(double), hipMemcpyDeviceToHost));
    if (vMF) AFQ_HIP(hipMemcpy(vMF, H->vMF, 2 * (size_t)H->K * sizeof(double), hipMemcpyDeviceToHost));
    if (P1) AFQ_HIP(hipMemcpy(P1, H->P1, 2 * MM * sizeof(double), hipMemcpyDeviceToHost));
    if (H1) AFQ_HIP(hipMemcpy(H1, H->H1, 2 * MM * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int afq_ham_stats(afq_ham H, long long* eloc3, long long* weight2, long long* popctl2) {
  return guard([&] {
    check_ham(H);
    unsigned long long st[7];
    AFQ_HIP(hipDeviceSynchronize());
    AFQ_HIP(hipMemcpy(st, H->stats, sizeof(st), hipMemcpyDeviceToHost));
    if (eloc3)
      for (int i = 0; i < 3; ++i) eloc3[i] = (long long)st[i];
    if (weight2)
      for (int i = 0; i < 2; ++i) weight2[i] = (long long)st[3 + i];
    if (popctl2)
      for (int i = 0; i < 2; ++i) popctl2[i] = (long long)st[5 + i];
  });
}

// ------------------------------------------------------------------ autotuner
}  // extern "C"

namespace {

// the cache: a JSON object of "key": tiling entries (string keys, integer values)
std::map<std::string, int> read_tile_cache(const char* path) {
  std::map<std::string, int> out;
  if (path == nullptr) return out;
  std::ifstream f(path);
  std::stringstream buf;
  buf << f.rdbuf();
  const std::string text = buf.str();
  size_t pos = 0;
  while (true) {
    const size_t a = text.find('"', pos);
    if (a == std::string::npos) break;
    const size_t b = text.find('"', a + 1);
    if (b == std::string::npos) break;
    const size_t c = text.find(':', b);
    if (c == std::string::npos) break;
    size_t e = c + 1;
    while (e < text.size() && (text[e] == ' ' || text[e] == '\t' || text[e] == '\n')) ++e;
    size_t f2 = e;
    while (f2 < text.size() && (std::isdigit((unsigned char)text[f2]) || text[f2] == '-')) ++f2;
    if (f2 > e) {
      try {
        out[text.substr(a + 1, b - a - 1)] = std::stoi(text.substr(e, f2 - e));
      } catch (...) {
      }
    }
    pos = f2 > b ? f2 : b + 1;
  }
  return out;
}

void write_tile_cache(const char* path, const std::map<std::string, int>& m) {
  if (path == nullptr) return;
  std::ofstream f(path);
  AFQ_CHECK(f.good(), AFQ_EINVAL, std::string("afq_autotune: cannot write ") + path);
  f << "{\n";
  size_t i = 0;
  for (const auto& kv : m) f << "  \"" << kv.first << "\": " << kv.second << (++i < m.size() ? ",\n" : "\n");
  f << "}\n";
}

// mean time (ms) of `reps` launches of fn on stream s
template <typename F>
float time_launches(hipStream_t s, int reps, F&& fn) {
  hipEvent_t a, b;
  AFQ_HIP(hipEventCreate(&a));
  AFQ_HIP(hipEventCreate(&b));
  float ms = 0.0f;
  try {
    fn();
    fn();
    AFQ_HIP(hipEventRecord(a, s));
    for (int r = 0; r < reps; ++r) fn();
    AFQ_HIP(hipEventRecord(b, s));
    AFQ_HIP(hipEventSynchronize(b));
    AFQ_HIP(hipEventElapsedTime(&ms, a, b));
  } catch (...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

}  // namespace

extern "C" {

int afq_autotune(afq_ham H, int W, const char* cache_path, int* cfg_out, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(W > 0 && W <= H->ws.cap, AFQ_ECAPACITY,
              "afq_autotune: W must be positive and within afq_ham_reserve's capacity");
    hipStream_t s = S(stream);
    const int M = H->M, N = H->N, K = H->K, p = H->precision;
    const long NM = (long)N * M;
    Workspace& ws = H->ws;
    auto cache = read_tile_cache(cache_path);
    auto key = [&](const char* what, long a, long b) {
      std::ostringstream k;
      k << what << "/p" << p << "/W" << W << "/" << a << "x" << b;
      return k.str();
    };
    auto tune = [&](const std::string& k, bool bc, auto&& launch) {
      const auto it = cache.find(k);
      if (it != cache.end() && it->second >= 0 && it->second < gemm_tile_configs(p, bc))
        return it->second;
      int best = 0;
      float tbest = 0.0f;
      for (int c = 0; c < gemm_tile_configs(p, bc); ++c) {
        const float t = time_launches(s, 10, [&] { launch(c); });
        if (c == 0 || t < tbest) {
          best = c;
          tbest = t;
        }
      }
      cache[k] = best;
      return best;
    };
    H->cfg_vbias = tune(key("vbias", K, NM), true,
                        [&](int c) { run_vbias(H, s, W, ws.G, ws.X, c); });
    H->cfg_vhs = tune(key("vhs", (long)M * M, K), false,
                      [&](int c) { run_vhs(H, s, W, ws.X, ws.V, c); });
    AFQ_HIP(hipStreamSynchronize(s));
    write_tile_cache(cache_path, cache);
    if (cfg_out) {
      cfg_out[0] = H->cfg_vbias;
      cfg_out[1] = H->cfg_vhs;
    }
  });
}

int afq_population_control(afq_ham H, int W, double* wbuf, long ldw, double uniform,
                           void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    AFQ_CHECK(uniform >= 0.0 && uniform < 1.0, AFQ_EINVAL, "uniform must be in [0, 1)");
    AFQ_CHECK(H->nfields == 0 || W <= H->fields_W, AFQ_ECAPACITY,
              "more walkers than the back-propagation field history holds");
    Workspace& ws = H->ws;
    BackPropRows bp;
    if (H->nfields > 0) {
      bp.fields = H->fields;
      bp.fields_ld = (long)H->nfields * H->K;
      bp.wfac = H->wfac;
      bp.whis = H->whis;
      bp.nhist = 3 * H->nfields;
      bp.pos = H->hist_pos;
    }
    launch_population_control(S(stream), W, wbuf, ldw, (long)H->M * H->N, uniform, ws.cum,
                              ws.off, ws.src, ws.ok, H->stats, H->nfields > 0 ? &bp : nullptr);
  });
}

// ------------------------------------------------------------------ det_ops level
int afq_det_mixed_density_matrix(int nb, int M, int N, const double* PsiT, const double* Phi,
                                 long ldw, double* G, double* ov, void* stream) {
  return guard([&] {
    AFQ_CHECK(nb >= 0 && M > 0 && N >= 0 && N <= M, AFQ_EBADSHAPE, "bad shape");
    AFQ_CHECK(ldw >= (long)M * N || nb <= 1, AFQ_EBADSHAPE, "walker stride smaller than M*N");
    AFQ_CHECK(ov != nullptr, AFQ_EINVAL, "null overlap array");
    launch_green(S(stream), nb, M, N, PsiT, Phi, ldw, G, ov, 2, 1.0, cmk(0, 0), nullptr, 0.0,
                 0.0, nullptr, 0);
  });
}

int afq_det_propagate(int nb, int M, int N, double* Phi, long ldw, const double* P1,
                      const double* V, int order, void* stream) {
  return guard([&] {
    AFQ_CHECK(nb >= 0 && M > 0 && N >= 0 && N <= M, AFQ_EBADSHAPE, "bad shape");
    AFQ_CHECK(order >= 0, AFQ_EINVAL, "negative Taylor order");
    launch_propagate(S(stream), nb, M, N, Phi, ldw, P1, V, order);
  });
}

int afq_det_orthogonalize(int nb, int M, int N, double* Phi, long ldw, double* logdetR,
                          void* stream) {
  return guard([&] {
    AFQ_CHECK(nb >= 0 && M > 0 && N >= 0 && N <= M, AFQ_EBADSHAPE, "bad shape");
    launch_orthogonalize(S(stream), nb, M, N, Phi, ldw, logdetR);
  });
}

// ------------------------------------------------------------------ wavefunction level
int afq_mixed_density_matrix(afq_ham H, int W, const double* wbuf, long ldw, double* G,
                             double* ov, void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    launch_green(S(stream), W, H->M, H->N, H->PsiT, wbuf, ldw, G, ov, ov ? 1 : 0, scl(H),
                 cmk(0, 0), nullptr, 0.0, 0.0, nullptr, 0);
  });
}

int afq_log_overlap(afq_ham H, int W, const double* wbuf, long ldw, double* ov, void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    AFQ_CHECK(ov != nullptr, AFQ_EINVAL, "null overlap array");
    launch_green(S(stream), W, H->M, H->N, H->PsiT, wbuf, ldw, nullptr, ov, 1, scl(H),
                 ci_log(H), nullptr, 0.0, 0.0, nullptr, 0);
  });
}

int afq_vbias(afq_ham H, int W, const double* G, double* v, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(W >= 0, AFQ_EBADSHAPE, "negative walker count");
    AFQ_CHECK(W <= H->ws.cap, AFQ_ECAPACITY, "more walkers than reserved");
    run_vbias(H, S(stream), W, G, v, H->cfg_vbias);
  });
}

int afq_vhs(afq_ham H, int W, const double* X, double* V, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(W >= 0, AFQ_EBADSHAPE, "negative walker count");
    run_vhs(H, S(stream), W, X, V, H->cfg_vhs);
  });
}

int afq_energy(afq_ham H, int W, const double* G, double* E, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(W <= H->ws.cap, AFQ_ECAPACITY, "more walkers than reserved");
    launch_e1(S(stream), W, H->M, H->N, G, H->haj, H->E0, scl(H), E);
    run_exchange(H, S(stream), W, G);
    finalize_energy(H, S(stream), W, E);
  });
}

int afq_wfn_energy(afq_ham H, int W, const double* wbuf, long ldw, double* E, double* ov,
                   void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    energy_from_walkers(H, S(stream), W, wbuf, ldw, E, ov);
  });
}

// ------------------------------------------------------------------ propagator level
int afq_construct_X(afq_ham H, int W, double* X, const double* xi, double* HW, double* mf,
                    void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(W >= 0, AFQ_EBADSHAPE, "negative walker count");
    const afq_step_params& p = H->params;
    launch_construct_X(S(stream), W, H->K, p.vbias_bound,
                       p.use_cp_constraint || p.project_force_bias, p.free_projection, H->vMF,
                       xi, X, HW, mf);
  });
}

int afq_propagate(afq_ham H, int W, double* wbuf, long ldw, const double* V, void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    launch_propagate(S(stream), W, H->M, H->N, wbuf, ldw, H->P1, V, H->params.taylor_n);
  });
}

int afq_walker_update(afq_ham H, int W, double* wbuf, long ldw, const double* ov,
                      const double* E, const double* mf, const double* HW, double Eshift,
                      int hybrid, void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    launch_walker_update(S(stream), W, wbuf, ldw, (long)H->M * H->N, H->params, H->dt, Eshift,
                         H->E0, ov, E, mf, HW, hybrid, H->stats);
  });
}

int afq_step(afq_ham H, int W, double* wbuf, long ldw, const double* xi, double Eshift,
             void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    AFQ_CHECK(xi != nullptr, AFQ_EINVAL, "null fields");
    AFQ_CHECK(H->nfields == 0 || W <= H->fields_W, AFQ_ECAPACITY,
              "more walkers than the back-propagation field history holds");
    hipStream_t s = S(stream);
    const int M = H->M, N = H->N, K = H->K;
    const long NM = (long)N * M;
    const afq_step_params& p = H->params;
    Workspace& ws = H->ws;
    Timers& tm = H->timers;
    // 1. vbias = 2 sqrt(dt) G Lnak^T   (NOMSD::vbias)
    {
      RegionScope r(tm, s, R_G_FOR_VBIAS);
      launch_green(s, W, M, N, H->PsiT, wbuf, ldw, ws.G, nullptr, 0, 1.0, cmk(0, 0), nullptr,
                   0.0, 0.0, nullptr, 0);
    }
    {
      RegionScope r(tm, s, R_VBIAS);
      run_vbias(H, s, W, ws.G, ws.X, H->cfg_vbias);
    }
    // 2. X, HW, mean-field factor     (assemble_X)
    {
      RegionScope r(tm, s, R_ASSEMBLE_X);
      launch_construct_X(s, W, K, p.vbias_bound, p.use_cp_constraint || p.project_force_bias,
                         p.free_projection, H->vMF, xi, ws.X, ws.HW, ws.mf);
      // the fields of this step into the back-propagation ring (storeFields)
      if (H->nfields > 0)
        launch_store_fields(s, W, K, ws.X, H->fields, H->nfields, H->hist_pos);
    }
    // 3. vHS = sqrt(dt) L X           4. Phi <- P1 exp(iV) P1 Phi
    {
      RegionScope r(tm, s, R_VHS);
      run_vhs(H, s, W, ws.X, ws.V, H->cfg_vhs);
    }
    {
      RegionScope r(tm, s, R_PROPAGATE);
      launch_propagate(s, W, M, N, wbuf, ldw, H->P1, ws.V, p.taylor_n);
    }
    // 5. new overlap (hybrid) or local energy
    if (p.hybrid) {
      RegionScope r(tm, s, R_OVERLAP_G_E1);
      launch_green(s, W, M, N, H->PsiT, wbuf, ldw, nullptr, ws.ov, 1, scl(H), ci_log(H),
                   nullptr, 0.0, 0.0, nullptr, 0);
    } else {
      {
        RegionScope r(tm, s, R_OVERLAP_G_E1);
        launch_green(s, W, M, N, H->PsiT, wbuf, ldw, ws.G, ws.ov, 1, scl(H), ci_log(H),
                     H->haj, H->E0, scl(H), ws.E, 3);
      }
      RegionScope r(tm, s, R_EXCHANGE);
      run_exchange(H, s, W, ws.G);
    }
    // 6.-7. energy assembly, weights, phases, energies, bounds
    {
      RegionScope r(tm, s, R_WALKER_UPDATE);
      if (!p.hybrid) finalize_energy(H, s, W, ws.E);
      launch_walker_update(s, W, wbuf, ldw, (long)M * N, p, H->dt, Eshift, H->E0, ws.ov, ws.E,
                           ws.mf, ws.HW, p.hybrid, H->stats, H->wfac, H->whis,
                           3 * H->nfields, H->hist_pos);
    }
    if (H->nfields > 0)   // advanceHistoryPos: wraps at the history length 3 nfields
      launch_advance_history(s, H->hist_pos, 3 * H->nfields);
  });
}

int afq_set_field_history(afq_ham H, int W, int nfields, double* d_fields, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(nfields >= 0 && W >= 0, AFQ_EBADSHAPE, "afq_set_field_history: negative size");
    AFQ_CHECK(nfields == 0 || d_fields != nullptr, AFQ_EINVAL,
              "afq_set_field_history: null field ring");
    if (!H->hist_pos) dalloc(&H->hist_pos, 1);
    // back-propagation's own device memory is sized here, so that neither the step nor
    // afq_back_propagate allocates
    if (nfields > 0) {
      if (!H->P1H) dalloc(&H->P1H, 2L * H->M * H->M);
      if (H->bp_cap < W) {
        dfree(H->bp_ldet);
        dalloc(&H->bp_ldet, 2L * std::max(W, 1));
        H->bp_cap = W;
      }
    }
    AFQ_HIP(hipMemsetAsync(H->hist_pos, 0, sizeof(int), S(stream)));
    AFQ_HIP(hipStreamSynchronize(S(stream)));
    H->fields = nfields > 0 ? d_fields : nullptr;
    H->nfields = nfields;
    H->fields_W = nfields > 0 ? W : 0;
    H->wfac = H->whis = nullptr;   // the weight rings belong to the previous field ring
  });
}

int afq_set_weight_history(afq_ham H, double* d_factors, double* d_history, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(H->nfields > 0, AFQ_EINVAL,
              "afq_set_weight_history: register the field history first");
    AFQ_CHECK((d_factors == nullptr) == (d_history == nullptr), AFQ_EINVAL,
              "afq_set_weight_history: give both rings or neither");
    if (d_factors != nullptr) {
      // WalkerSetBase::resize_bp: factors start at 1, the weight history at 0
      const long n = (long)H->fields_W * 3 * H->nfields;
      launch_ph_fill(S(stream), n, 1.0, 0.0, d_factors);
      launch_ph_fill(S(stream), n, 0.0, 0.0, d_history);
      AFQ_HIP(hipStreamSynchronize(S(stream)));
    }
    H->wfac = d_factors;
    H->whis = d_history;
  });
}

int afq_history_position(afq_ham H, int* pos, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(pos != nullptr, AFQ_EINVAL, "afq_history_position: null output");
    *pos = 0;
    if (!H->hist_pos) return;
    AFQ_HIP(hipMemcpyAsync(pos, H->hist_pos, sizeof(int), hipMemcpyDeviceToHost, S(stream)));
    AFQ_HIP(hipStreamSynchronize(S(stream)));
  });
}

int afq_back_propagate(afq_ham H, int W, int nbpsteps, int nstabilize, double* d_refs, int nrefs,
                       double* d_logdetR, void* stream) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(H->nfields > 0, AFQ_EINVAL, "afq_back_propagate: no field history registered");
    AFQ_CHECK(nbpsteps > 0, AFQ_EINVAL, "afq_back_propagate: nbpsteps must be positive");
    AFQ_CHECK(nbpsteps <= H->nfields, AFQ_EBADSHAPE,
              "afq_back_propagate: more steps than the field history holds");
    AFQ_CHECK(nstabilize > 0, AFQ_EINVAL, "afq_back_propagate: nstabilize must be positive");
    AFQ_CHECK(W >= 0 && W <= H->fields_W && W <= H->ws.cap, AFQ_ECAPACITY,
              "afq_back_propagate: more walkers than the field history or workspace holds");
    AFQ_CHECK(nrefs >= 1, AFQ_EBADSHAPE, "afq_back_propagate: no references");
    AFQ_CHECK(d_refs && d_logdetR, AFQ_EINVAL, "afq_back_propagate: null operand");
    if (W == 0) return;
    hipStream_t s = S(stream);
    const int M = H->M, N = H->N, K = H->K, MM = M * M, nf = H->nfields;
    const long NM = (long)N * M;
    AFQ_CHECK(H->P1H != nullptr && H->bp_cap >= W, AFQ_ECAPACITY,
              "afq_back_propagate: more walkers than afq_set_field_history sized it for");
    int pos = 0;
    AFQ_HIP(hipMemcpyAsync(&pos, H->hist_pos, sizeof(int), hipMemcpyDeviceToHost, s));
    AFQ_HIP(hipStreamSynchronize(s));
    launch_adjoint(s, M, H->P1, H->P1H);
    launch_ph_fill(s, (long)W * nrefs, 0.0, 0.0, d_logdetR);
    int slot = pos % nf;
    for (int ni = nbpsteps - 1; ni >= 0; --ni) {
      // newest entry first: the position already points at the next step's slot
      slot = slot == 0 ? nf - 1 : slot - 1;
      // vHS = sqrt(dt) L X of the stored fields (FP64), then V <- -V^H
      launch_gemm_nt_z(s, AFQ_FP64, W, MM, K, std::sqrt(H->dt), H->fields + 2L * slot * K,
                       (long)nf * K, H->Likn, K, false, H->ws.V, MM);
      launch_neg_adjoint(s, W, M, H->ws.V);
      for (int nr = 0; nr < nrefs; ++nr) {
        double* Phi = d_refs + 2L * nr * NM;
        const long ld = (long)nrefs * NM;
        launch_propagate(s, W, M, N, Phi, ld, H->P1H, H->ws.V, H->params.taylor_n);
        if (ni == 0 || ni % nstabilize == 0) {
          launch_ph_fill(s, W, 0.0, 0.0, H->bp_ldet);
          launch_orthogonalize(s, W, M, N, Phi, ld, H->bp_ldet);
          launch_add_strided(s, W, H->bp_ldet, d_logdetR + 2L * nr, nrefs);
        }
      }
    }
  });
}

int afq_orthogonalize(afq_ham H, int W, double* wbuf, long ldw, void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    hipStream_t s = S(stream);
    RegionScope r(H->timers, s, R_ORTHO);
    launch_orthogonalize(s, W, H->M, H->N, wbuf, ldw, nullptr);
    launch_green(s, W, H->M, H->N, H->PsiT, wbuf, ldw, nullptr, H->ws.ov, 1, scl(H),
                 ci_log(H), nullptr, 0.0, 0.0, nullptr, 0);
    launch_set_properties(s, W, wbuf, ldw, (long)H->M * H->N, H->ws.ov, nullptr, 0);
  });
}

int afq_seed_properties(afq_ham H, int W, double* wbuf, long ldw, void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    hipStream_t s = S(stream);
    energy_from_walkers(H, s, W, wbuf, ldw, H->ws.E, H->ws.ov);
    launch_set_properties(s, W, wbuf, ldw, (long)H->M * H->N, H->ws.ov, H->ws.E, 1);
  });
}

int afq_weighted_sums(afq_ham H, int W, const double* wbuf, long ldw, double* out,
                      void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    launch_weighted_sums(S(stream), W, wbuf, ldw, (long)H->M * H->N, H->dt, out);
  });
}

int afq_estimator_sums(afq_ham H, int W, const double* wbuf, long ldw, const double* E,
                       const double* ov, double* out, void* stream) {
  return guard([&] {
    check_ham(H);
    check_walkers(H, W, ldw);
    launch_estimator_sums(S(stream), W, wbuf, ldw, (long)H->M * H->N, H->dt, E, ov, out);
  });
}

int afq_ham_enable_timers(afq_ham H, int max_records) {
  return guard([&] {
    check_ham(H);
    AFQ_CHECK(max_records >= 0, AFQ_EINVAL, "negative record count");
    AFQ_HIP(hipDeviceSynchronize());
    destroy_timers(H->timers);
    for (int r = 0; r < R_COUNT; ++r) {
      H->timers.ms[r] = 0.0;
      H->timers.count[r] = 0;
    }
    if (max_records == 0) return;
    H->timers.ev.resize(2 * (size_t)max_records);
    for (auto& e : H->timers.ev) AFQ_HIP(hipEventCreate(&e));
    H->timers.region.assign(max_records, 0);
    H->timers.cap = max_records;
    H->timers.on = true;
  });
}

int afq_ham_read_timers(afq_ham H, double* h_ms, long long* h_counts) {
  return guard([&] {
    check_ham(H);
    Timers& t = H->timers;
    for (int i = 0; i < t.used; ++i) {
      AFQ_HIP(hipEventSynchronize(t.ev[2 * i + 1]));
      float ms = 0.f;
      AFQ_HIP(hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]));
      t.ms[t.region[i]] += ms;
      t.count[t.region[i]] += 1;
    }
    t.used = 0;
    for (int r = 0; r < R_COUNT; ++r) {
      if (h_ms) h_ms[r] = t.ms[r];
      if (h_counts) h_counts[r] = t.count[r];
      t.ms[r] = 0.0;
      t.count[r] = 0;
    }
  });
}

const char* afq_timer_name(int region) {
  return (region >= 0 && region < R_COUNT) ? kRegionNames[region] : "";
}

int afq_ph_excited_2body_energy(int nw, int ndet, int nex, int nact, int nelec, int nchol,
                                const int* iexcit, const int* refc, const double* Twina,
                                const double* R, const double* wgt, double* EX, double* EJ,
                                double* KE, void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && ndet >= 0 && nact > 0 && nelec > 0 && nelec <= nact && nchol >= 0,
              AFQ_EBADSHAPE, "ph 2-body energy: bad sizes");
    // the staged kernel's group partials: a temporary of this call (the PHMSD handle keeps its
    // own for the walker step)
    const long need = ph_2body_workspace(nw, nact, nelec, nchol);
    double* ws = nullptr;
    if (need > 0 && nw > 0 && ndet > 0) AFQ_HIP(hipMalloc(&ws, need * sizeof(double)));
    try {
      launch_ph_2body(S(stream), nw, ndet, nex, nact, nelec, nchol, iexcit, refc, Twina, R,
                      wgt, EX, EJ, KE, ws, need);
      if (ws) AFQ_HIP(hipStreamSynchronize(S(stream)));
    } catch (...) {
      if (ws) (void)hipFree(ws);
      throw;
    }
    if (ws) AFQ_HIP(hipFree(ws));
  });
}

}  // extern "C"

// Scratch memory of afq_ph_excited_2body_energy_noalloc: the group partials of the staged
// 2-body kernel, 4 doubles per (walker, group of Cholesky vectors). Each device holds its own
// copy for the life of the process. Host code never passes `scratch` itself (its host-side
// name is not a device address): the device address comes from hipGetSymbolAddress, once per
// device, and only that address reaches kernels.
static constexpr long kNoAllocScratchDoubles = 1L << 20;   // 8 MiB per device
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, resolved on the first call on that device.
double* device_scratch() {
  int dev = 0;
  AFQ_HIP(hipGetDevice(&dev));
  AFQ_CHECK(dev >= 0 && dev < kMaxDevices, AFQ_EUNSUPPORTED, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    AFQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AFQ_CHECK(p != nullptr, AFQ_EHIP, "hipGetSymbolAddress returned no address for scratch");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

long afq_ph_noalloc_scratch_doubles(void) { return kNoAllocScratchDoubles; }

int afq_ph_excited_2body_energy_noalloc(int nw, int ndet, int nex, int nact, int nelec,
                                        int nchol, const int* iexcit, const int* refc,
                                        const double* Twina, const double* R, const double* wgt,
                                        double* EX, double* EJ, double* KE, void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && ndet >= 0 && nact > 0 && nelec > 0 && nelec <= nact && nchol >= 0,
              AFQ_EBADSHAPE, "ph 2-body energy: bad sizes");
    const long need = ph_2body_workspace(nw, nact, nelec, nchol);
    AFQ_CHECK(need <= kNoAllocScratchDoubles, AFQ_ECAPACITY,
              "ph 2-body energy (no allocation): the group partials need " +
                  std::to_string(need) + " doubles, more than the " +
                  std::to_string(kNoAllocScratchDoubles) +
                  " of the static scratch buffer (4 * walkers * ceil(nchol / 4) <= 2^20)");
    // every group partial is written before k_ph_2body_finish reads it: nothing to clear
    double* ws = (need > 0 && nw > 0 && ndet > 0) ? device_scratch() : nullptr;
    launch_ph_2body(S(stream), nw, ndet, nex, nact, nelec, nchol, iexcit, refc, Twina, R, wgt,
                    EX, EJ, KE, ws, ws ? kNoAllocScratchDoubles : 0);
  });
}

int afq_ph_excited_1body_energy(int nw, int ndet, int nex, int nact, int nelec,
                                const int* iexcit, const int* refc, const double* Smat,
                                const double* R, const double* wgt, double* E, void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && ndet >= 0 && nex >= 0 && nact > 0 && nelec > 0 && nelec <= nact,
              AFQ_EBADSHAPE, "ph 1-body energy: bad sizes");
    launch_ph_1body(S(stream), nw, ndet, nex, nact, nelec, iexcit, refc, Smat, R, wgt, E);
  });
}

int afq_ph_mixed_density_matrix_woodbury(int nb, int nmo, int nel, int nact, const double* A,
                                         const int* ref, const double* Phi, long ldw, double* G,
                                         double* ov, double* QQ0, double* work, void* stream) {
  return guard([&] {
    AFQ_CHECK(nb >= 0 && nmo > 0 && nel > 0 && nel <= nact && nact <= nmo &&
                  ldw >= (long)nmo * nel,
              AFQ_EBADSHAPE, "Woodbury mixed density matrix: bad sizes");
    AFQ_CHECK(nb == 0 || (A && ref && Phi && G && ov && QQ0 && work), AFQ_EINVAL,
              "Woodbury mixed density matrix: null operand");
    launch_ph_woodbury(S(stream), nb, nmo, nel, nact, A, ref, Phi, ldw, G, ov, QQ0, work);
  });
}

int afq_ph_reference_energy(int nw, int nel, int nmo, int nchol, int nact, double scl, double E0,
                            const double* G, const double* Lna, const double* haj, double* E,
                            double* Kl, double* Twina, double* Swia, void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && nel > 0 && nmo > 0 && nchol > 0 && nact >= nel, AFQ_EBADSHAPE,
              "ph reference energy: bad sizes (needs nact >= nel)");
    AFQ_CHECK(G && Lna && E && Kl && Twina && (!haj || Swia), AFQ_EINVAL,
              "ph reference energy: null operand");
    launch_ph_reference_energy(S(stream), nw, nel, nmo, nchol, nact, scl, E0, G, Lna, haj, E, Kl,
                               Twina, Swia);
  });
}

int afq_ph_spin_weights(int nw, int nu, int nd, const int* cp, const int* ci, const double* cv,
                        const double* rup, const double* rdn, double* wgt, double* ov,
                        void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && nu >= 0 && nd >= 0, AFQ_EBADSHAPE, "ph spin weights: bad sizes");
    launch_ph_spin_weights(S(stream), nw, nu, cp, ci, cv, rup, rdn, wgt, ov);
  });
}

int afq_ph_overlaps(int nw, int ndet, int nex, int nact, int nelec, const int* iexcit,
                    const double* T, double* ov, void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && ndet >= 0 && nex >= 1 && nact > 0 && nelec > 0 && nelec <= nact,
              AFQ_EBADSHAPE, "ph overlaps: bad sizes");
    launch_ph_overlaps(S(stream), nw, ndet, nex, nact, nelec, iexcit, T, ov);
  });
}

int afq_ph_compact_R(int nw, int ndet, int nex, int nact, int nelec, const int* iexcit,
                     const int* refc, const double* T, double* R, void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && ndet >= 0 && nex >= 1 && nact > 0 && nelec > 0 && nelec <= nact,
              AFQ_EBADSHAPE, "ph R matrices: bad sizes");
    launch_ph_compact_R(S(stream), nw, ndet, nex, nact, nelec, iexcit, refc, T, R);
  });
}

int afq_ph_accumulate_R(int nw, int ndet, int nex, int nact, int nelec, const int* iexcit,
                        const int* refc, const double* Rc, const double* ov, const double* wgt,
                        double* R, void* stream) {
  return guard([&] {
    AFQ_CHECK(nw >= 0 && ndet >= 0 && nex >= 0 && nact > 0 && nelec > 0 && nelec <= nact,
              AFQ_EBADSHAPE, "ph R accumulation: bad sizes");
    launch_ph_accumulate_R(S(stream), nw, ndet, nex, nact, nelec, iexcit, refc, Rc, ov, wgt, R);
  });
}

int afq_probe_mfma(int kind, double* out, void* stream) {
  return guard([&] {
    AFQ_CHECK(kind >= 0 && kind <= 5, AFQ_EINVAL,
              "probe kind must be 0 (FP64), 1 (FP32), 2 or 3 (BF16), 4 or 5 (int8)");
    launch_probe_mfma(S(stream), kind, out);
  });
}

}  // extern "C"
