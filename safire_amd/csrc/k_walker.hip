// Per-walker scalar work of a step, kept on the GPU (the reference stages the walker properties
// to the host every step, WalkerSetUpdate.hpp:89-100, 174-183):
//   k_construct_X    construct_X_impl (construct_X_kernel.hpp:33-84) + meanfield factor
//                    (AFQMCBasePropagator.cpp:334-336); one workgroup per walker, fixed-order sums
//   k_walker_update  walker_update + bound_local_energy + bound_walker_weights
//                    (WalkerSetUpdate.hpp:46-218); one thread per walker
//   k_set_properties OVLP/E1/EXX/EJ/PSEUDO_ELOC seeding (NOMSD.cpp:69-80, execute.cpp:172-184)
//   k_weighted_sums  the sums of averagePseudoEnergy and averageEnergy
//   k_estimator_sums the sums of EnergyEstimator::measure and weight_statistics
#include "afq_internal.hpp"

#include <cfloat>

namespace afq {

namespace {

__device__ inline cd clamp_bound(cd z, double bound) {
  const double a = cabs_(z);
  if (a > bound) return (bound / a) * z;   // z/|z| * bound
  return z;
}

__global__ __launch_bounds__(256) void k_construct_X(int W, int K, double bound, int project,
                                                     int free_projection,
                                                     const cd* __restrict__ vMF,
                                                     const double* __restrict__ xi,
                                                     cd* __restrict__ X, cd* __restrict__ HW,
                                                     cd* __restrict__ mf) {
  __shared__ double red[4][6];
  const int w = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cd hw = cmk(0, 0), xv = cmk(0, 0);
  double v2 = 0.0;
  for (int m = tid; m < K; m += 256) {
    const cd vmf0 = vMF[m];
    const cd vmf_t = clamp_bound(vmf0, bound);
    cd vb_t = clamp_bound(X[(long)w * K + m], bound);
    if (project) vb_t = cmk(vb_t.re, 0.0);
    const cd d = vb_t - vmf_t;
    const cd vdiff = free_projection ? cmk(0, 0) : cmk(-d.im, d.re);   // i (vb - vmf)
    const cd x = cmk(xi[(long)w * K + m], 0.0) + vdiff;
    X[(long)w * K + m] = x;
    hw -= vdiff * (x - 0.5 * vdiff);
    xv += x * vmf0;
    v2 += (vmf0 * vmf0).re;
  }
  double v[6] = {hw.re, hw.im, xv.re, xv.im, v2, 0.0};
#pragma unroll
  for (int q = 0; q < 5; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 5; ++q) red[wave][q] = v[q];
  __syncthreads();
  if (tid == 0) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int qw = 0; qw < 4; ++qw)
      for (int q = 0; q < 5; ++q) s[q] += red[qw][q];
    HW[w] = cmk(s[0], s[1]);
    if (mf) mf[w] = cmk(-0.5 * s[4] - s[3], s[2]);   // -Re(sum vMF^2)/2 + i sum X vMF
  }
}

__global__ void k_walker_update(int W, cd* __restrict__ wbuf, long ldw, long poff, int hybrid,
                                int free_projection, int use_cp, double dt, double Eshift,
                                double E0, const cd* __restrict__ new_ov,
                                const cd* __restrict__ E, const cd* __restrict__ mf,
                                const cd* __restrict__ HW, double lower_scale,
                                double upper_scale, double max_weight,
                                unsigned long long* __restrict__ stats, cd* __restrict__ wfac,
                                cd* __restrict__ whis, int nhist, const int* __restrict__ pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  cd factor = cmk(0.0, 0.0);   // back-propagation's weight factor (0 for a dropped walker)
  cd* p = wbuf + (long)i * ldw + poff;
  const cd old_ov = p[AFQ_P_OVLP], old_eloc = p[AFQ_P_PSEUDO_ELOC];
  double scale = 1.0, dtheta = 0.0;
  if (!free_projection) {
    dtheta = new_ov[i].im - old_ov.im - mf[i].im;
    const double c = cos(dtheta);
    scale = use_cp ? (c > 0.0 ? 1.0 : 0.0) : fmax(0.0, c);
  }
  p[AFQ_P_THETA] = cmk(dtheta, 0.0);
  cd eloc;
  if (hybrid) {
    const cd num = mf[i] - HW[i] - (new_ov[i] - old_ov);
    eloc = cmk(num.re / dt + E0, num.im / dt);
  } else {
    eloc = E[3 * i] + E[3 * i + 1] + E[3 * i + 2];
  }
  unsigned long long n_upper = 0, n_lower = 0;
  if (!isfinite(eloc.re)) {
    p[AFQ_P_WEIGHT] = cmk(0, 0);
    p[AFQ_P_PHASE] = cmk(0, 0);
    eloc = old_eloc;
  } else {
    if (!free_projection) {
      const double hi = Eshift + upper_scale * sqrt(2.0 / dt);
      const double lo = Eshift - lower_scale * sqrt(2.0 / dt);
      if (eloc.re > hi) n_upper = 1;
      else if (eloc.re < lo) n_lower = 1;
      eloc.re = fmin(fmax(eloc.re, lo), hi);
    }
    const cd eeff = hybrid ? eloc : 0.5 * (eloc + old_eloc);
    const cd wf = cexp_(cmk(-dt * (eeff.re - Eshift), -dt * eeff.im));
    const double aw = cabs_(wf);
    const cd wt = p[AFQ_P_WEIGHT];
    p[AFQ_P_WEIGHT] = free_projection ? (wt * (scale * wf)) : (scale * aw) * wt;
    p[AFQ_P_PHASE] = p[AFQ_P_PHASE] * cmk(wf.re / aw, wf.im / aw);
    if (fabs(scale) > 2.2250738585072014e-308) factor = cmk(wf.re / aw / scale, wf.im / aw / scale);
  }
  // with back-propagation the factor and the updated weight (before the bound below) go into
  // the history rings at the current position (WalkerSetUpdate.hpp:184-190)
  if (wfac != nullptr) {
    const long slot = (long)i * nhist + *pos;
    wfac[slot] = factor;
    whis[slot] = p[AFQ_P_WEIGHT];
  }
  p[AFQ_P_PSEUDO_ELOC] = eloc;
  p[AFQ_P_OVLP] = new_ov[i];
  if (!hybrid) {
    p[AFQ_P_E1] = E[3 * i];
    p[AFQ_P_EXX] = E[3 * i + 1];
    p[AFQ_P_EJ] = E[3 * i + 2];
  }
  // bound_walker_weights: |w| capped at max(floor, fraction * target population)
  unsigned long long n_wcap = 0;
  const cd wt = p[AFQ_P_WEIGHT];
  const double a = cabs_(wt);
  if (a > max_weight) {
    n_wcap = 1;
    p[AFQ_P_WEIGHT] = (max_weight / a) * wt;
  }
  if (stats) {
    if (!free_projection && isfinite(eloc.re)) atomicAdd(&stats[0], 1ull);
    if (n_upper) atomicAdd(&stats[1], 1ull);
    if (n_lower) atomicAdd(&stats[2], 1ull);
    atomicAdd(&stats[3], 1ull);
    if (n_wcap) atomicAdd(&stats[4], 1ull);
  }
}

__global__ void k_set_properties(int W, cd* __restrict__ wbuf, long ldw, long poff,
                                 const cd* __restrict__ ov, const cd* __restrict__ E,
                                 int seed_pseudo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  cd* p = wbuf + (long)i * ldw + poff;
  p[AFQ_P_OVLP] = ov[i];
  if (E) {
    p[AFQ_P_E1] = E[3 * i];
    p[AFQ_P_EXX] = E[3 * i + 1];
    p[AFQ_P_EJ] = E[3 * i + 2];
    if (seed_pseudo) p[AFQ_P_PSEUDO_ELOC] = E[3 * i] + E[3 * i + 1] + E[3 * i + 2];
  }
}

// out[0] = sum w (E1+EXX+EJ), out[1] = sum w, out[2] = sum w exp(dt PSEUDO_ELOC),
// out[3] = sum w PSEUDO_ELOC; one workgroup, fixed order
__global__ __launch_bounds__(256) void k_weighted_sums(int W, const cd* __restrict__ wbuf,
                                                       long ldw, long poff, double dt,
                                                       cd* __restrict__ out) {
  __shared__ cd red[4][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cd s[4] = {cmk(0, 0), cmk(0, 0), cmk(0, 0), cmk(0, 0)};
  for (int i = tid; i < W; i += 256) {
    const cd* p = wbuf + (long)i * ldw + poff;
    const cd w = p[AFQ_P_WEIGHT];
    const cd e = p[AFQ_P_E1] + p[AFQ_P_EXX] + p[AFQ_P_EJ];
    const cd pe = p[AFQ_P_PSEUDO_ELOC];
    s[0] += w * e;
    s[1] += w;
    s[2] += w * cexp_(dt * pe);
    s[3] += w * pe;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    for (int off = 32; off > 0; off >>= 1) {
      s[q].re += __shfl_xor(s[q].re, off, 64);
      s[q].im += __shfl_xor(s[q].im, off, 64);
    }
  if (lane == 0)
    for (int q = 0; q < 4; ++q) red[wave][q] = s[q];
  __syncthreads();
  if (tid < 4) {
    cd t = cmk(0, 0);
    for (int qw = 0; qw < 4; ++qw) t += red[qw][tid];
    out[tid] = t;
  }
}

// The walker sums of one estimator measurement, in a fixed order (one workgroup):
//   EnergyEstimator::measure (EnergyEstimator.h:51-100): sum_w E_w,c w_w for c = E1, EXX, EJ
//     (gemv of the transposed energies with the weights, no conjugation), sum_w w_w, and the
//     overlap sum_w exp(OVLP_w) w_w; E and OVLP come from the walker slots (walkers carrying
//     their energy) or from E [W][3] / ov [W] when given;
//   weight_statistics (Measurements.hpp:130-149): sum |w|, sum |w|^2;
//   averagePseudoEnergy / averageEnergy (average_energy.hpp:53-79): sum w exp(dt PSEUDO_ELOC),
//     sum w (E1 + EXX + EJ) from the walker slots.
__global__ __launch_bounds__(256) void k_estimator_sums(int W, const cd* __restrict__ wbuf,
                                                        long ldw, long poff, double dt,
                                                        const cd* __restrict__ E,
                                                        const cd* __restrict__ ov,
                                                        cd* __restrict__ out) {
  constexpr int NS = 9;
  __shared__ cd red[4][NS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cd s[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) s[q] = cmk(0, 0);
  for (int i = tid; i < W; i += 256) {
    const cd* p = wbuf + (long)i * ldw + poff;
    const cd w = p[AFQ_P_WEIGHT];
    const cd e1 = E ? E[3L * i] : p[AFQ_P_E1];
    const cd exx = E ? E[3L * i + 1] : p[AFQ_P_EXX];
    const cd ej = E ? E[3L * i + 2] : p[AFQ_P_EJ];
    const cd o = ov ? ov[i] : p[AFQ_P_OVLP];
    const double aw = cabs_(w);
    s[0] += e1 * w;
    s[1] += exx * w;
    s[2] += ej * w;
    s[3] += w;
    s[4] += cexp_(o) * w;
    s[5].re += aw;
    s[6].re += aw * aw;
    s[7] += w * cexp_(dt * p[AFQ_P_PSEUDO_ELOC]);
    s[8] += w * (p[AFQ_P_E1] + p[AFQ_P_EXX] + p[AFQ_P_EJ]);
  }
#pragma unroll
  for (int q = 0; q < NS; ++q)
    for (int off = 32; off > 0; off >>= 1) {
      s[q].re += __shfl_xor(s[q].re, off, 64);
      s[q].im += __shfl_xor(s[q].im, off, 64);
    }
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NS; ++q) red[wave][q] = s[q];
  __syncthreads();
  if (tid < NS) {
    cd t = cmk(0, 0);
    for (int qw = 0; qw < 4; ++qw) t += red[qw][tid];
    out[tid] = t;
  }
}

// Population control by the comb (population_control.hpp:34-65) and single-rank placement
// (swap_walkers_async, population_control.hpp:143-209). The comb is a sequential walk over the
// cumulative weights; one thread runs it in the reference's order so that the offspring are
// bit-identical. src[slot] = walker copied into slot, or -1 when the slot keeps its walker.
__global__ void k_comb_plan(int W, const cd* __restrict__ wbuf, long ldw, long poff, double r,
                            double* __restrict__ cum, int* __restrict__ off,
                            int* __restrict__ src, int* __restrict__ ok,
                            unsigned long long* __restrict__ stats) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ok[0] = 0;
  double total = 0.0;
  int last = -1;
  for (int i = 0; i < W; ++i) {
    const double a = cabs_(wbuf[(long)i * ldw + poff + AFQ_P_WEIGHT]);
    total += a;
    cum[i] = total;
    off[i] = 0;
    src[i] = -1;
    if (a > 0.0) last = i;
  }
  if (stats) atomicAdd(&stats[5], 1ull);
  if (!(total > 1e-6) || last < 0) {
    if (stats) atomicAdd(&stats[6], 1ull);   // collapsed: walkers left unchanged
    return;
  }
  int iw = 0, ic = 0;
  while (ic < W) {
    const double comb = (ic + r) * total / W;
    if (comb < cum[iw] || iw == last) {
      off[iw] += 1;
      ic++;
    } else {
      iw++;
    }
  }
  int next = 0;   // next zero-offspring slot
  for (int i = 0; i < W; ++i)
    for (int c = 1; c < off[i]; ++c) {
      while (off[next] != 0) ++next;
      src[next] = i;
      ++next;
    }
  ok[0] = 1;
}

// The same comb in parallel, one workgroup. The cumulative weights are summed in the
// reference's sequential order (one thread, from LDS), because their rounding decides ties.
// Tooth ic sits at comb_ic = (ic + r) * total / W (the reference's expression, monotone in ic);
// it lands on the first walker i with comb_ic < cum[i], or on the last walker with weight.
// So offspring[i] = count(cum[i]) - count(cum[i-1]) with count(x) = #{ic : comb_ic < x}, found
// by binary search; the copy destinations follow from exclusive scans of the extra copies and of
// the zero-offspring slots (integers, exact).
__device__ inline double comb_tooth(int ic, double r, double total, int W) {
  return ((double)ic + r) * total / W;
}

__device__ inline int comb_count(double x, double r, double total, int W) {
  int lo = 0, hi = W;   // first ic with comb_ic >= x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (comb_tooth(mid, r, total, W) < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// exclusive scan of v[0..n) in place (integers), one workgroup of 256 threads
__device__ void block_exclusive_scan(int* v, int n, int* tmp256) {
  const int t = threadIdx.x;
  const int per = (n + 255) / 256;
  const int b = t * per, e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; ++i) s += v[i];
  tmp256[t] = s;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int q = 0; q < 256; ++q) {
      const int x = tmp256[q];
      tmp256[q] = acc;
      acc += x;
    }
  }
  __syncthreads();
  int acc = tmp256[t];
  for (int i = b; i < e; ++i) {
    const int x = v[i];
    v[i] = acc;
    acc += x;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_comb_parallel(int W, const cd* __restrict__ wbuf,
                                                       long ldw, long poff, double r,
                                                       int* __restrict__ src,
                                                       int* __restrict__ ok,
                                                       unsigned long long* __restrict__ stats) {
  extern __shared__ double smem[];
  double* cum = smem;                               // [W]
  int* off = reinterpret_cast<int*>(cum + W);       // [W]
  int* ext = off + W;                               // [W] extra copies -> base
  int* zer = ext + W;                               // [W] zero flags -> free index
  __shared__ int tmp[256];
  __shared__ int s_last;
  __shared__ double s_total;
  const int t = threadIdx.x;
  for (int i = t; i < W; i += 256) {
    cum[i] = cabs_(wbuf[(long)i * ldw + poff + AFQ_P_WEIGHT]);
    src[i] = -1;
  }
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    int last = -1;
    for (int i = 0; i < W; ++i) {
      const double a = cum[i];
      if (a > 0.0) last = i;
      s += a;
      cum[i] = s;
    }
    s_total = s;
    s_last = last;
    ok[0] = 0;
    if (stats) atomicAdd(&stats[5], 1ull);
    if (!(s > 1e-6) || last < 0) {
      if (stats) atomicAdd(&stats[6], 1ull);   // collapsed: walkers left unchanged
    }
  }
  __syncthreads();
  const double total = s_total;
  const int last = s_last;
  if (!(total > 1e-6) || last < 0) return;
  for (int i = t; i < W; i += 256) {
    int n = 0;
    if (i < last) {
      const int hi = comb_count(cum[i], r, total, W);
      const int lo = i > 0 ? comb_count(cum[i - 1], r, total, W) : 0;
      n = hi - lo;
    } else if (i == last) {
      n = W - (i > 0 ? comb_count(cum[i - 1], r, total, W) : 0);
    }
    off[i] = n;
    ext[i] = n > 1 ? n - 1 : 0;
    zer[i] = n == 0 ? 1 : 0;
  }
  __syncthreads();
  block_exclusive_scan(ext, W, tmp);   // ext[i]: first free-list entry of walker i's copies
  block_exclusive_scan(zer, W, tmp);   // zer[i]: free-list entry of zero-offspring slot i
  int* freelist = reinterpret_cast<int*>(cum);      // cum is no longer needed
  for (int i = t; i < W; i += 256)
    if (off[i] == 0) freelist[zer[i]] = i;
  __syncthreads();
  for (int i = t; i < W; i += 256) {
    const int n = off[i];
    for (int c = 1; c < n; ++c) src[freelist[ext[i] + c - 1]] = i;
  }
  if (t == 0) ok[0] = 1;
}

__global__ void k_copy_walkers(int W, cd* __restrict__ wbuf, long ldw,
                               const int* __restrict__ src) {
  const int slot = blockIdx.x;
  const int from = src[slot];
  if (from < 0) return;
  for (long e = threadIdx.x; e < ldw; e += blockDim.x)
    wbuf[(long)slot * ldw + e] = wbuf[(long)from * ldw + e];
}

__global__ void k_normalize_weights(int W, cd* __restrict__ wbuf, long ldw, long poff,
                                    const int* __restrict__ ok) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W) return;
  if (ok[0] == 0) return;   // the comb refused a collapsed population: weights unchanged
  cd& w = wbuf[(long)i * ldw + poff + AFQ_P_WEIGHT];
  const double a = cabs_(w);
  w = cmk(w.re / a, w.im / a);
}

// D = A B with A[i][k] = i + 100 k, B[k][j] = 1 if k == j%4 else 0  =>  D[i][j] = i + 100 (j%4);
// out[lane][r] holds accumulator element r of each lane, for reading back the lane layout.
template <typename R>
__global__ void k_probe_mfma(double* out) {
  typedef typename acc_of<R>::type acc_t;
  const int lane = threadIdx.x;
  const int i = lane & 15, k = lane >> 4, j = lane & 15;
  const R a = (R)(i + 100 * k);
  const R b = (k == (j & 3)) ? (R)1 : (R)0;
  acc_t acc = {0, 0, 0, 0};
  acc = mfma16x16x4(a, b, acc);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = (double)acc[r];
}

// ---- back-propagation (AFQMCBasePropagator.cpp:336-344, 387-390, 403-532)

// WalkerSet::storeFields: ring[w][pos % nfields][n] = X[w][n] (the history position lives on the
// GPU so that a captured step replays it)
__global__ void k_store_fields(int W, int K, const cd* __restrict__ X, cd* __restrict__ ring,
                               int nfields, const int* __restrict__ pos) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)W * K) return;
  const int slot = *pos % nfields;
  const long w = e / K, n = e % K;
  ring[(w * nfields + slot) * K + n] = X[e];
}

// WalkerSet::advanceHistoryPos: the position wraps at the history length, 3 x the field ring
// (WalkerSetBase.h:158, WalkerSetBase.cpp:340)
__global__ void k_advance_history(int* pos, int nhist) {
  if (threadIdx.x == 0) *pos = (*pos + 1) % nhist;
}

// rows[slot] <- rows[src[slot]] for the slots population control refills (src >= 0); a source
// keeps its own slot, so no row is both read and written
__global__ void k_copy_rows(int W, cd* __restrict__ rows, long ld, long ncols,
                            const int* __restrict__ src) {
  const int slot = blockIdx.x;
  const int from = src[slot];
  if (from < 0) return;
  for (long e = threadIdx.x; e < ncols; e += blockDim.x)
    rows[(long)slot * ld + e] = rows[(long)from * ld + e];
}

// after the comb the last completed step's weight-history entry holds the normalised weight
// (population_control.hpp:249-254)
__global__ void k_history_fixup(int W, const cd* __restrict__ wbuf, long ldw, long poff,
                                cd* __restrict__ whis, int nhist, const int* __restrict__ pos,
                                const int* __restrict__ ok) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W || ok[0] == 0) return;
  const int slot = (*pos + nhist - 1) % nhist;
  whis[(long)i * nhist + slot] = wbuf[(long)i * ldw + poff + AFQ_P_WEIGHT];
}

// V_w <- -V_w^H in place, so that the forward propagator's exp(i V) series becomes the 'H'
// series exp(-i V^H) of apply_expM<'H'> (propagate.hpp:57-58, 96-97). One thread per pair i <= j.
__global__ void k_neg_adjoint(int M, cd* __restrict__ V) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y;
  if (e >= (long)M * M) return;
  const int i = (int)(e / M), j = (int)(e % M);
  if (i > j) return;
  cd* Vw = V + (long)w * M * M;
  const cd a = Vw[(long)i * M + j], b = Vw[(long)j * M + i];
  Vw[(long)i * M + j] = cmk(-b.re, b.im);
  Vw[(long)j * M + i] = cmk(-a.re, a.im);
}

// B = A^H for one [M][M] matrix (math::product<'H'>(P1, ...), propagate.hpp:282, 297)
__global__ void k_adjoint(int M, const cd* __restrict__ A, cd* __restrict__ B) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)M * M) return;
  const int i = (int)(e / M), j = (int)(e % M);
  const cd a = A[(long)j * M + i];
  B[e] = cmk(a.re, -a.im);
}

// out[w * stride] += x[w]
__global__ void k_add_strided(int W, const cd* __restrict__ x, cd* __restrict__ out,
                              long stride) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < W) out[(long)w * stride] += x[w];
}


}  // namespace

void launch_construct_X(hipStream_t s, int W, int K, double vbias_bound, int project,
                        int free_projection, const double* vMF, const double* xi, double* X,
                        double* HW, double* mf) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_construct_X, dim3(W), dim3(256), 0, s, W, K, vbias_bound, project,
                     free_projection, reinterpret_cast<const cd*>(vMF), xi,
                     reinterpret_cast<cd*>(X), reinterpret_cast<cd*>(HW),
                     reinterpret_cast<cd*>(mf));
  AFQ_HIP(hipGetLastError());
}

void launch_walker_update(hipStream_t s, int W, double* wbuf, long ldw, long poff,
                          const afq_step_params& p, double dt, double Eshift, double E0,
                          const double* ov, const double* E, const double* mf,
                          const double* HW, int hybrid, unsigned long long* stats,
                          double* wfac, double* whis, int nhist, const int* pos) {
  if (W <= 0) return;
  const double maxw = fmax(p.weight_bound_floor, p.weight_bound_fraction * (double)W);
  const bool hist = wfac != nullptr && whis != nullptr && nhist > 0 && pos != nullptr;
  hipLaunchKernelGGL(k_walker_update, dim3((W + 127) / 128), dim3(128), 0, s, W,
                     reinterpret_cast<cd*>(wbuf), ldw, poff, hybrid, p.free_projection,
                     p.use_cp_constraint, dt, Eshift, E0, reinterpret_cast<const cd*>(ov),
                     reinterpret_cast<const cd*>(E), reinterpret_cast<const cd*>(mf),
                     reinterpret_cast<const cd*>(HW), p.lower_cutoff_scale,
                     p.upper_cutoff_scale, maxw, stats,
                     hist ? reinterpret_cast<cd*>(wfac) : nullptr,
                     hist ? reinterpret_cast<cd*>(whis) : nullptr, hist ? nhist : 0,
                     hist ? pos : nullptr);
  AFQ_HIP(hipGetLastError());
}

void launch_set_properties(hipStream_t s, int W, double* wbuf, long ldw, long poff,
                           const double* ov, const double* E, int seed_pseudo) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_set_properties, dim3((W + 127) / 128), dim3(128), 0, s, W,
                     reinterpret_cast<cd*>(wbuf), ldw, poff, reinterpret_cast<const cd*>(ov),
                     reinterpret_cast<const cd*>(E), seed_pseudo);
  AFQ_HIP(hipGetLastError());
}

void launch_weighted_sums(hipStream_t s, int W, const double* wbuf, long ldw, long poff,
                          double dt, double* out) {
  hipLaunchKernelGGL(k_weighted_sums, dim3(1), dim3(256), 0, s, W,
                     reinterpret_cast<const cd*>(wbuf), ldw, poff, dt,
                     reinterpret_cast<cd*>(out));
  AFQ_HIP(hipGetLastError());
}

void launch_estimator_sums(hipStream_t s, int W, const double* wbuf, long ldw, long poff,
                           double dt, const double* E, const double* ov, double* out) {
  hipLaunchKernelGGL(k_estimator_sums, dim3(1), dim3(256), 0, s, W,
                     reinterpret_cast<const cd*>(wbuf), ldw, poff, dt,
                     reinterpret_cast<const cd*>(E), reinterpret_cast<const cd*>(ov),
                     reinterpret_cast<cd*>(out));
  AFQ_HIP(hipGetLastError());
}

void launch_population_control(hipStream_t s, int W, double* wbuf, long ldw, long poff,
                               double r, double* cum, int* off, int* src, int* ok,
                               unsigned long long* stats, const BackPropRows* bp) {
  if (W <= 0) return;
  const size_t lds = (size_t)W * (sizeof(double) + 3 * sizeof(int));
  if (lds <= 120 * 1024) {
    ensure_dynamic_lds((const void*)k_comb_parallel, lds);
    hipLaunchKernelGGL(k_comb_parallel, dim3(1), dim3(256), lds, s, W,
                       reinterpret_cast<const cd*>(wbuf), ldw, poff, r, src, ok, stats);
  } else {
    hipLaunchKernelGGL(k_comb_plan, dim3(1), dim3(64), 0, s, W,
                       reinterpret_cast<const cd*>(wbuf), ldw, poff, r, cum, off, src, ok, stats);
  }
  hipLaunchKernelGGL(k_copy_walkers, dim3(W), dim3(256), 0, s, W, reinterpret_cast<cd*>(wbuf),
                     ldw, (const int*)src);
  if (bp != nullptr && bp->fields != nullptr) {
    // WalkerSet::copy_walker moves the back-propagation row with the walker
    // (WalkerSetBase.h:377-382, 413-415): the field ring, and the weight rings when present
    hipLaunchKernelGGL(k_copy_rows, dim3(W), dim3(256), 0, s, W,
                       reinterpret_cast<cd*>(bp->fields), bp->fields_ld, bp->fields_ld,
                       (const int*)src);
    if (bp->wfac != nullptr) {
      hipLaunchKernelGGL(k_copy_rows, dim3(W), dim3(64), 0, s, W, reinterpret_cast<cd*>(bp->wfac),
                         (long)bp->nhist, (long)bp->nhist, (const int*)src);
      hipLaunchKernelGGL(k_copy_rows, dim3(W), dim3(64), 0, s, W, reinterpret_cast<cd*>(bp->whis),
                         (long)bp->nhist, (long)bp->nhist, (const int*)src);
    }
  }
  hipLaunchKernelGGL(k_normalize_weights, dim3((W + 127) / 128), dim3(128), 0, s, W,
                     reinterpret_cast<cd*>(wbuf), ldw, poff, (const int*)ok);
  if (bp != nullptr && bp->whis != nullptr)
    hipLaunchKernelGGL(k_history_fixup, dim3((W + 127) / 128), dim3(128), 0, s, W,
                       reinterpret_cast<const cd*>(wbuf), ldw, poff,
                       reinterpret_cast<cd*>(bp->whis), bp->nhist, bp->pos, (const int*)ok);
  AFQ_HIP(hipGetLastError());
}

// BF16 16x16x32: A[i][k] = (i + 1) and B[k][j] = (j + 1) for k == ks only, so D[i][j] =
// (i + 1)(j + 1). Lane l supplies A[l%16][8*(l/16) + q] and B[8*(l/16) + q][l%16], q = 0..7.
__global__ void k_probe_mfma_bf16(int ks, double* out) {
  const int lane = threadIdx.x;
  const int i = lane & 15, g = lane >> 4;
  bf16x8 a, b;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = 8 * g + q;
    a[q] = (__bf16)(k == ks ? (float)(i + 1) : 0.0f);
    b[q] = (__bf16)(k == ks ? (float)(i + 1) : 0.0f);   // column j = lane % 16 = i
  }
  f4 acc = {0, 0, 0, 0};
  acc = mfma_bf16(a, b, acc);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = (double)acc[r];
}

// int8 16x16x64: lane l supplies bytes 16 (l / 16) + j, j < 16, of row / column l % 16
__global__ void k_probe_mfma_i8(int ks, double* out) {
  typedef int i4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x;
  const int i = lane & 15, g = lane >> 4;
  i4 a = {0, 0, 0, 0};
  if (ks / 16 == g) a[(ks % 16) / 4] = (i + 1) << (8 * (ks % 4));
  i4 acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, a, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = (double)acc[r];
}

void launch_probe_mfma(hipStream_t s, int kind, double* out) {
  if (kind == 0)
    hipLaunchKernelGGL(k_probe_mfma<double>, dim3(1), dim3(64), 0, s, out);
  else if (kind == 1)
    hipLaunchKernelGGL(k_probe_mfma<float>, dim3(1), dim3(64), 0, s, out);
  else if (kind <= 3)
    hipLaunchKernelGGL(k_probe_mfma_bf16, dim3(1), dim3(64), 0, s, kind == 2 ? 0 : 31, out);
  else
    hipLaunchKernelGGL(k_probe_mfma_i8, dim3(1), dim3(64), 0, s, kind == 4 ? 0 : 63, out);
  AFQ_HIP(hipGetLastError());
}

void launch_store_fields(hipStream_t s, int W, int K, const double* X, double* ring, int nfields,
                         const int* pos) {
  if (W <= 0 || K <= 0) return;
  hipLaunchKernelGGL(k_store_fields, dim3((unsigned)(((long)W * K + 255) / 256)), dim3(256), 0, s,
                     W, K, reinterpret_cast<const cd*>(X), reinterpret_cast<cd*>(ring), nfields,
                     pos);
  AFQ_HIP(hipGetLastError());
}

void launch_advance_history(hipStream_t s, int* pos, int nhist) {
  hipLaunchKernelGGL(k_advance_history, dim3(1), dim3(64), 0, s, pos, nhist);
  AFQ_HIP(hipGetLastError());
}

void launch_neg_adjoint(hipStream_t s, int W, int M, double* V) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_neg_adjoint, dim3((unsigned)(((long)M * M + 255) / 256), W), dim3(256), 0,
                     s, M, reinterpret_cast<cd*>(V));
  AFQ_HIP(hipGetLastError());
}

void launch_adjoint(hipStream_t s, int M, const double* A, double* B) {
  hipLaunchKernelGGL(k_adjoint, dim3((unsigned)(((long)M * M + 255) / 256)), dim3(256), 0, s, M,
                     reinterpret_cast<const cd*>(A), reinterpret_cast<cd*>(B));
  AFQ_HIP(hipGetLastError());
}

void launch_add_strided(hipStream_t s, int W, const double* x, double* out, long stride) {
  if (W <= 0) return;
  hipLaunchKernelGGL(k_add_strided, dim3((unsigned)((W + 255) / 256)), dim3(256), 0, s, W,
                     reinterpret_cast<const cd*>(x), reinterpret_cast<cd*>(out), stride);
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
