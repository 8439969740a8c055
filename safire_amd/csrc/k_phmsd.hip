// Particle-hole multi-determinant (PHMSD) energy corrections of the excited determinants,
// detail/phmsd_impl.hpp:
//   ph_excited_2body_energy_dense_cholesky (29-147): EX[w] -= sum_d wgt[d][w] eX_d / 2,
//     EJ[w] += sum_d wgt[d][w] (sum_n KE[d][w][n]^2 - eJ0_w) / 2, KE[d][w][n] written
//   ph_excited_1body_energy (149-183): E[w] += sum_d wgt[d][w] e_d
// for excitations iexcit[d] = (electron rows p_1..p_nex of refc, then their target orbitals),
// the walker tensor Twina[w][i][n][a] (electron i, Cholesky vector n, active orbital a) and the
// compact R[w][d][e][a] of each determinant.
//
// One workgroup per walker walks the determinants in order; its threads split the Cholesky index
// n, so every term of eX (all of them are sums over n) and KE[d][w][n] come out per thread
// without a reduction per determinant. The weight of a determinant is a scalar, so each thread
// keeps weighted running sums and the workgroup reduces them once, in a fixed order (no atomics:
// results do not depend on scheduling). R[w][d] and the determinant's occupation are staged in
// LDS per determinant. Every index read from the excitation lists is clamped into range before
// it addresses Twina; a determinant with an out-of-range index makes the walker's results NaN.
#include "afq_internal.hpp"

#include <cfloat>

namespace afq {

namespace {

constexpr int kPhThreads = 256;
constexpr int kPhMaxEx = 8;        // excitation rank handled per determinant
constexpr int kPhMaxAct = 256;     // active orbitals
constexpr int kPhMaxElec = 64;
constexpr long kP2Lds = 64 * 1024;   // LDS bound of the staged 2-body kernel (bytes)

__device__ inline cd block_sum(cd v, cd* red) {
  const int t = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) {
    v.re += __shfl_xor(v.re, off, 64);
    v.im += __shfl_xor(v.im, off, 64);
  }
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  cd s = cmk(0.0, 0.0);
  for (int i = 0; i < kPhThreads / 64; ++i) s += red[i];   // fixed order
  __syncthreads();
  return s;
}

template <int NEX>
__global__ __launch_bounds__(kPhThreads) void k_ph_2body(
    int nw, int ndet, int nact, int nelec, int nchol, const int* __restrict__ iexcit,
    const int* __restrict__ refc, const cd* __restrict__ Twina, const cd* __restrict__ R,
    const cd* __restrict__ wgt, cd* __restrict__ EX, cd* __restrict__ EJ, cd* __restrict__ KE) {
  __shared__ cd sR[NEX * kPhMaxAct];
  __shared__ int sOcc[kPhMaxElec], sRef[kPhMaxElec], sP[NEX];
  __shared__ int sBad;
  __shared__ cd red[kPhThreads / 64];
  const int w = blockIdx.x;
  const int t = threadIdx.x;
  const cd* Tw = Twina + (long)w * nelec * nchol * nact;
  auto T = [&](int i, int n, int a) { return Tw[((long)i * nchol + n) * nact + a]; };
  if (t == 0) {
    int bad = 0;
    for (int i = 0; i < nelec; ++i) bad |= (refc[i] < 0 || refc[i] >= nact);
    sBad = bad;
  }
  for (int i = t; i < nelec; i += kPhThreads) sRef[i] = min(max(refc[i], 0), nact - 1);
  __syncthreads();
  const bool bad_ref = sBad != 0;

  // eJ0 = sum_n (sum_i T[i][n][refc i])^2, as thread partials over n
  cd ej0 = cmk(0.0, 0.0);
  for (int n = t; n < nchol; n += kPhThreads) {
    cd f0 = cmk(0.0, 0.0);
    for (int i = 0; i < nelec; ++i) f0 += T(i, n, sRef[i]);
    ej0 += f0 * f0;
  }

  cd ex_acc = cmk(0.0, 0.0), ej_acc = cmk(0.0, 0.0);
  for (int d = 0; d < ndet; ++d) {
    const int* ex = iexcit + (long)d * 2 * NEX;
    __syncthreads();   // the previous determinant's LDS reads are done
    for (int e = t; e < NEX * nact; e += kPhThreads)
      sR[e] = R[(((long)w * ndet + d) * NEX) * nact + e];
    if (t < NEX) sP[t] = min(max(ex[t], 0), nelec - 1);
    if (t == 0) {
      int bad = bad_ref ? 1 : 0;
      for (int e = 0; e < NEX; ++e)
        bad |= (ex[e] < 0 || ex[e] >= nelec || ex[NEX + e] < 0 || ex[NEX + e] >= nact);
      sBad = bad;
    }
    for (int i = t; i < nelec; i += kPhThreads) {
      int o = sRef[i];
      for (int e = 0; e < NEX; ++e)
        if (ex[e] == i) {
          o = min(max(ex[NEX + e], 0), nact - 1);
          break;
        }
      sOcc[i] = o;
    }
    __syncthreads();

    cd eX = cmk(0.0, 0.0), kk = cmk(0.0, 0.0);
    for (int n = t; n < nchol; n += kPhThreads) {
      // Y_ef = sum_a T[p_e][n][a] R_f[a]: the diagonal terms and the symmetric pairs
      cd kn = cmk(0.0, 0.0);
#pragma unroll
      for (int e = 0; e < NEX; ++e) {
        cd yee = cmk(0.0, 0.0);
        for (int a = 0; a < nact; ++a) yee += T(sP[e], n, a) * sR[e * nact + a];
        eX += yee * yee;
        kn += yee;
#pragma unroll
        for (int f = e + 1; f < NEX; ++f) {
          cd yfe = cmk(0.0, 0.0), yef = cmk(0.0, 0.0);
          for (int a = 0; a < nact; ++a) {
            yfe += T(sP[f], n, a) * sR[e * nact + a];
            yef += T(sP[e], n, a) * sR[f * nact + a];
          }
          eX += 2.0 * (yfe * yef);
        }
        // R_e against the occupied columns: sum_j T[p_e][n][occ j] (T[j][n][:] . R_e)
        for (int j = 0; j < nelec; ++j) {
          cd z = cmk(0.0, 0.0);
          for (int a = 0; a < nact; ++a) z += T(j, n, a) * sR[e * nact + a];
          eX += 2.0 * (T(sP[e], n, sOcc[j]) * z);
        }
      }
      // rows whose occupation changed
      for (int i = 0; i < nelec; ++i) {
        const int oi = sOcc[i], ri = sRef[i];
        const cd tio = T(i, n, oi), tir = T(i, n, ri);
        if (oi != ri) eX += tio * tio - tir * tir;
        for (int j = i + 1; j < nelec; ++j) {
          const int oj = sOcc[j], rj = sRef[j];
          if (oi != ri || oj != rj)
            eX += 2.0 * (T(i, n, oj) * T(j, n, oi) - T(i, n, rj) * T(j, n, ri));
        }
        kn += tio;
      }
      KE[((long)d * nw + w) * nchol + n] = kn;
      kk += kn * kn;
    }
    const cd wd = wgt[(long)d * nw + w];
    if (sBad) eX = cmk(__builtin_nan(""), __builtin_nan(""));
    ex_acc += wd * eX;
    ej_acc += wd * (kk - ej0);
  }
  const cd sx = block_sum(ex_acc, red);
  const cd sj = block_sum(ej_acc, red);
  if (t == 0) {
    EX[w] -= 0.5 * sx;
    EJ[w] += 0.5 * sj;
  }
}

// E[w] += sum_d wgt[d][w] sum_e (S[w][p_e] . R[w][d][e] + S[w][p_e][v_e] - S[w][p_e][refc p_e])
__global__ __launch_bounds__(kPhThreads) void k_ph_1body(
    int nw, int ndet, int nex, int nact, int nelec, const int* __restrict__ iexcit,
    const int* __restrict__ refc, const cd* __restrict__ S, const cd* __restrict__ R,
    const cd* __restrict__ wgt, cd* __restrict__ E) {
  __shared__ cd red[kPhThreads / 64];
  const int w = blockIdx.x;
  cd acc = cmk(0.0, 0.0);
  for (int d = threadIdx.x; d < ndet; d += kPhThreads) {
    const int* ex = iexcit + (long)d * 2 * nex;
    cd e_ = cmk(0.0, 0.0);
    for (int e = 0; e < nex; ++e) {
      const bool ok = ex[e] >= 0 && ex[e] < nelec && ex[nex + e] >= 0 && ex[nex + e] < nact;
      const int p = ok ? ex[e] : 0, v = ok ? ex[nex + e] : 0;
      const int rp = (ok && refc[p] >= 0 && refc[p] < nact) ? refc[p] : -1;
      const cd* Sp = S + ((long)w * nelec + p) * nact;
      const cd* Re = R + (((long)w * ndet + d) * nex + e) * nact;
      for (int a = 0; a < nact; ++a) e_ += Sp[a] * Re[a];
      if (rp < 0) e_ = cmk(__builtin_nan(""), __builtin_nan(""));
      else e_ += Sp[v] - Sp[rp];
    }
    acc += wgt[(long)d * nw + w] * e_;
  }
  const cd s = block_sum(acc, red);
  if (threadIdx.x == 0) E[w] += s;
}

// ---- excitation overlaps and R matrices (phmsd_helpers.hpp) ----
//
// For the rank-nex excitations of one spin, iexcit [ndet][2 nex] holds the excited electrons
// e_0..e_{nex-1} and their target active orbitals e_nex..e_{2nex-1}; T [nw][nact][nel] is the
// walker's active-space matrix. Per determinant and walker A(p, q) = T[e_{nex+p}][e_q]:
//   overlap ratio = det A                                   (calculate_overlaps, 111-195)
//   compact R(p, a), Q = A^-1, zero unless a is occupied in the excited configuration orbs:
//     R(p, orbs(i)) = sum_q Q(p, q) (delta(i, e_q) - T[e_{nex+q}][i])
//                                                          (get_compact_ph_R_matrices, 317-383)
//   and nothing where det A == 0.
// One thread per (determinant, walker); A is factorised in registers by Gauss-Jordan with
// partial pivoting (first maximum of |re| + |im|).
template <int NEX>
__device__ inline bool small_inverse(cd (&a)[NEX][NEX], cd (&q)[NEX][NEX], cd& det) {
#pragma unroll
  for (int i = 0; i < NEX; ++i)
#pragma unroll
    for (int j = 0; j < NEX; ++j) q[i][j] = cmk(i == j ? 1.0 : 0.0, 0.0);
  det = cmk(1.0, 0.0);
#pragma unroll
  for (int j = 0; j < NEX; ++j) {
    int p = j;
    double best = cabs1(a[j][j]);
#pragma unroll
    for (int r = j + 1; r < NEX; ++r)
      if (cabs1(a[r][j]) > best) {
        best = cabs1(a[r][j]);
        p = r;
      }
    if (p != j) {
      det = -det;
#pragma unroll
      for (int r = j + 1; r < NEX; ++r)
        if (r == p)
#pragma unroll
          for (int c = 0; c < NEX; ++c) {
            const cd t = a[j][c];
            a[j][c] = a[r][c];
            a[r][c] = t;
            const cd u = q[j][c];
            q[j][c] = q[r][c];
            q[r][c] = u;
          }
    }
    const cd piv = a[j][j];
    det = det * piv;
    if (cis0(piv)) return false;
    const cd inv = cdiv(cmk(1.0, 0.0), piv);
#pragma unroll
    for (int c = 0; c < NEX; ++c) {
      a[j][c] = a[j][c] * inv;
      q[j][c] = q[j][c] * inv;
    }
#pragma unroll
    for (int r = 0; r < NEX; ++r) {
      if (r == j) continue;
      const cd f = a[r][j];
#pragma unroll
      for (int c = 0; c < NEX; ++c) {
        a[r][c] -= f * a[j][c];
        q[r][c] -= f * q[j][c];
      }
    }
  }
  return true;
}

template <int NEX>
__device__ inline bool load_excitation(int nact, int nelec, const int* e, const cd* Tw,
                                       cd (&a)[NEX][NEX]) {
  bool ok = true;
#pragma unroll
  for (int p = 0; p < NEX; ++p) ok &= e[p] >= 0 && e[p] < nelec && e[NEX + p] >= 0 &&
                                      e[NEX + p] < nact;
#pragma unroll
  for (int p = 0; p < NEX; ++p)
#pragma unroll
    for (int q = 0; q < NEX; ++q) {
      const int r = min(max(e[NEX + p], 0), nact - 1), c = min(max(e[q], 0), nelec - 1);
      a[p][q] = Tw[(long)r * nelec + c];
    }
  return ok;
}

template <int NEX>
__global__ __launch_bounds__(256) void k_ph_overlaps(int nw, int ndet, int nact, int nelec,
                                                     const int* __restrict__ iexcit,
                                                     const cd* __restrict__ T,
                                                     cd* __restrict__ ov) {
  const long id = (long)blockIdx.x * 256 + threadIdx.x;
  if (id >= (long)nw * ndet) return;
  const int w = (int)(id % nw), d = (int)(id / nw);
  cd a[NEX][NEX], q[NEX][NEX], det;
  const bool ok = load_excitation<NEX>(nact, nelec, iexcit + (long)d * 2 * NEX,
                                       T + (long)w * nact * nelec, a);
  if (!small_inverse<NEX>(a, q, det)) det = cmk(0.0, 0.0);
  ov[(long)d * nw + w] = ok ? det : cmk(__builtin_nan(""), __builtin_nan(""));
}

template <int NEX>
__global__ __launch_bounds__(256) void k_ph_compact_R(int nw, int ndet, int nact, int nelec,
                                                      const int* __restrict__ iexcit,
                                                      const int* __restrict__ refc,
                                                      const cd* __restrict__ T,
                                                      cd* __restrict__ R) {
  const long id = (long)blockIdx.x * 256 + threadIdx.x;
  if (id >= (long)nw * ndet) return;
  const int w = (int)(id / ndet), d = (int)(id % ndet);
  const int* e = iexcit + (long)d * 2 * NEX;
  const cd* Tw = T + (long)w * nact * nelec;
  cd* Rd = R + ((long)w * ndet + d) * NEX * nact;
  for (int i = 0; i < NEX * nact; ++i) Rd[i] = cmk(0.0, 0.0);
  cd a[NEX][NEX], q[NEX][NEX], det;
  const bool ok = load_excitation<NEX>(nact, nelec, e, Tw, a);
  if (!ok) {
    for (int i = 0; i < NEX * nact; ++i) Rd[i] = cmk(__builtin_nan(""), __builtin_nan(""));
    return;
  }
  if (!small_inverse<NEX>(a, q, det)) return;
  for (int i = 0; i < nelec; ++i) {
    int o = min(max(refc[i], 0), nact - 1);
    int qi = -1;
#pragma unroll
    for (int p = 0; p < NEX; ++p)
      if (e[p] == i) {
        o = e[NEX + p];
        qi = p;
      }
#pragma unroll
    for (int p = 0; p < NEX; ++p) {
      cd v = cmk(0.0, 0.0);
#pragma unroll
      for (int qq = 0; qq < NEX; ++qq) v -= q[p][qq] * Tw[(long)e[NEX + qq] * nelec + i];
      if (qi >= 0) v += q[p][qi];
      Rd[(long)p * nact + o] += v;
    }
  }
}

// calculate_R (phmsd_helpers.hpp:200-262) for one shell of rank nex, added to R [nw][nel][nact]:
//   R[w][i][orbs_d(i)] += wgt[d][w] for every electron, and R[w][e_p][:] += wgt[d][w] Rc[w][d][p][:]
// for every determinant with det A != 0 (ov != 0), in determinant order; nex = 0 adds the
// reference determinant's term R[w][i][refc(i)] += wgt[0][w]. One workgroup per walker, threads
// over the entries (i, a), so every entry is summed over the determinants in their order.
__global__ __launch_bounds__(256) void k_ph_accumulate_R(int nw, int ndet, int nex, int nact,
                                                         int nelec, const int* __restrict__ iexcit,
                                                         const int* __restrict__ refc,
                                                         const cd* __restrict__ Rc,
                                                         const cd* __restrict__ ov,
                                                         const cd* __restrict__ wgt,
                                                         cd* __restrict__ R) {
  const int w = blockIdx.x;
  cd* Rw = R + (long)w * nelec * nact;
  for (int ea = threadIdx.x; ea < nelec * nact; ea += 256) {
    const int i = ea / nact, a = ea % nact;
    cd acc = Rw[ea];
    if (nex == 0) {
      if (refc[i] == a) acc += wgt[w];
    } else {
      for (int d = 0; d < ndet; ++d) {
        if (cis0(ov[(long)d * nw + w])) continue;
        const int* e = iexcit + (long)d * 2 * nex;
        const cd wd = wgt[(long)d * nw + w];
        int o = refc[i], pe = -1;
        for (int p = 0; p < nex; ++p)
          if (e[p] == i) {
            o = e[nex + p];
            pe = p;
          }
        if (o == a) acc += wd;
        if (pe >= 0) acc += wd * Rc[(((long)w * ndet + d) * nex + pe) * nact + a];
      }
    }
    Rw[ea] = acc;
  }
}

// MixedDensityMatrixForWoodbury (density_matrix.hpp:651-717) after the reference determinant's
// G: QQ0[w][a][e] = sum_k A[a][k] G[w][e][k] (= TAB inv(TNN), since G^T = Phi inv(TNN)); one
// thread per output entry.
__global__ __launch_bounds__(256) void k_ph_qq0(int nb, int nmo, int nel, int nact,
                                                const cd* __restrict__ A,
                                                const cd* __restrict__ G,
                                                cd* __restrict__ QQ0) {
  const long id = (long)blockIdx.x * 256 + threadIdx.x;
  if (id >= (long)nb * nact * nel) return;
  const int e = (int)(id % nel);
  const int a = (int)((id / nel) % nact);
  const long w = id / ((long)nel * nact);
  const cd* Ar = A + (long)a * nmo;
  const cd* Gr = G + (w * nel + e) * nmo;
  cd acc = cmk(0.0, 0.0);
  for (int k = 0; k < nmo; ++k) acc += Ar[k] * Gr[k];
  QQ0[id] = acc;
}

// rows ref[i] of A [nact][nmo] into Aref [nel][nmo]
__global__ __launch_bounds__(256) void k_gather_rows(int nel, int nmo, int nact,
                                                     const int* __restrict__ ref,
                                                     const cd* __restrict__ A,
                                                     cd* __restrict__ Aref) {
  const long id = (long)blockIdx.x * 256 + threadIdx.x;
  if (id >= (long)nel * nmo) return;
  const int i = (int)(id / nmo), k = (int)(id % nmo);
  const int r = ref[i];
  Aref[id] = (r >= 0 && r < nact) ? A[(long)r * nmo + k]
                                  : cmk(__builtin_nan(""), __builtin_nan(""));
}

// The spin coupling of the determinant weights (PHMSD.cpp:441-450): with the CI couplings
// C [nu][nd] in CSR form (row pointers cp, columns ci, values cv),
//   wgt[u][w] = ratio_up[u][w] sum_{d in row u} C[u][d] ratio_dn[d][w]   (csrmm, then the product)
//   Ov[w]     = sum_u wgt[u][w]                                          (contract "uw,uw->w")
// One workgroup per walker; each thread takes a fixed set of rows u and the workgroup sums the
// partial overlaps in a fixed order.
__global__ __launch_bounds__(256) void k_ph_spin_weights(int nw, int nu, const int* __restrict__ cp,
                                                         const int* __restrict__ ci,
                                                         const cd* __restrict__ cv,
                                                         const cd* __restrict__ rup,
                                                         const cd* __restrict__ rdn,
                                                         cd* __restrict__ wgt,
                                                         cd* __restrict__ ov) {
  __shared__ cd red[kPhThreads / 64];
  const int w = blockIdx.x;
  cd part = cmk(0.0, 0.0);
  for (int u = threadIdx.x; u < nu; u += kPhThreads) {
    cd acc = cmk(0.0, 0.0);
    for (int j = cp[u]; j < cp[u + 1]; ++j) acc += cv[j] * rdn[(long)ci[j] * nw + w];
    const cd x = acc * rup[(long)u * nw + w];
    wgt[(long)u * nw + w] = x;
    part += x;
  }
  const cd sum = block_sum(part, red);
  if (threadIdx.x == 0) ov[w] = sum;
}

// The same 2-body energy with the Cholesky index split over workgroups: a workgroup takes one
// walker and kP2N consecutive vectors n, stages T[i][n][a] of those vectors in LDS once and
// sweeps every determinant against them (kP2D determinant lanes per vector), so Twina is read
// from memory once per walker instead of once per determinant. Per (n, d) the thread forms
// Z[j][e] = sum_a T[j][n][a] R_e[a] row by row from LDS (R_d from the L1 cache) and keeps
// Y[e][f] = Z[p_e][f]; the terms are those of k_ph_2body. Each workgroup writes its weighted
// sums (a fixed-order workgroup reduction) to part[w][group]; k_ph_2body_finish adds the groups
// of a walker in order.
constexpr int kP2N = 4;     // Cholesky vectors per workgroup
constexpr int kP2D = 32;    // determinant lanes per vector
constexpr int kP2Threads = kP2N * kP2D;

template <int NEX>
__global__ __launch_bounds__(kP2Threads) void k_ph_2body_lds(
    int nw, int ndet, int nact, int nelec, int nchol, int ngrp, const int* __restrict__ iexcit,
    const int* __restrict__ refc, const cd* __restrict__ Twina, const cd* __restrict__ R,
    const cd* __restrict__ wgt, cd* __restrict__ part, cd* __restrict__ KE) {
  extern __shared__ cd sT[];                 // [kP2N][nelec][nact]
  __shared__ int sRef[kPhMaxElec];
  __shared__ int sBad;
  __shared__ cd red[kP2Threads / 64];
  const int w = blockIdx.x, grp = blockIdx.y;
  const int t = threadIdx.x, nl = t / kP2D, dl = t % kP2D;
  const int n0 = grp * kP2N, n = n0 + nl;
  const cd* Tw = Twina + (long)w * nelec * nchol * nact;
  const int per = nelec * nact;
  for (int e = t; e < kP2N * per; e += kP2Threads) {
    const int q = e / per, rem = e - q * per, i = rem / nact, a = rem - i * nact;
    sT[e] = (n0 + q < nchol) ? Tw[((long)i * nchol + n0 + q) * nact + a] : cmk(0.0, 0.0);
  }
  if (t == 0) {
    int bad = 0;
    for (int i = 0; i < nelec; ++i) bad |= (refc[i] < 0 || refc[i] >= nact);
    sBad = bad;
  }
  for (int i = t; i < nelec; i += kP2Threads) sRef[i] = min(max(refc[i], 0), nact - 1);
  __syncthreads();
  const bool bad_ref = sBad != 0;
  const cd* Tn = sT + (long)nl * per;        // T[i][n][a] = Tn[i * nact + a]
  cd ex_acc = cmk(0.0, 0.0), ej_acc = cmk(0.0, 0.0);
  if (n < nchol) {
    cd f0 = cmk(0.0, 0.0);
    for (int i = 0; i < nelec; ++i) f0 += Tn[i * nact + sRef[i]];
    const cd ej0 = f0 * f0;
    for (int d = dl; d < ndet; d += kP2D) {
      const int* ex = iexcit + (long)d * 2 * NEX;
      int hole[NEX], p[NEX], v[NEX];
      bool bad = bad_ref;
#pragma unroll
      for (int e = 0; e < NEX; ++e) {
        hole[e] = ex[e];
        bad |= (ex[e] < 0 || ex[e] >= nelec || ex[NEX + e] < 0 || ex[NEX + e] >= nact);
        p[e] = min(max(ex[e], 0), nelec - 1);
        v[e] = min(max(ex[NEX + e], 0), nact - 1);
      }
      // occupation of row i: the reference's, or the orbital the first excitation of i goes to
      auto occ = [&](int i) {
        int o = sRef[i];
        bool hit = false;
#pragma unroll
        for (int e = 0; e < NEX; ++e)
          if (!hit && hole[e] == i) {
            o = v[e];
            hit = true;
          }
        return o;
      };
      const cd* Rd = R + (((long)w * ndet + d) * NEX) * nact;
      cd Y[NEX][NEX];
#pragma unroll
      for (int e = 0; e < NEX; ++e)
#pragma unroll
        for (int f = 0; f < NEX; ++f) Y[e][f] = cmk(0.0, 0.0);
      cd eX = cmk(0.0, 0.0), kn = cmk(0.0, 0.0);
      for (int j = 0; j < nelec; ++j) {
        cd z[NEX];
#pragma unroll
        for (int e = 0; e < NEX; ++e) z[e] = cmk(0.0, 0.0);
        const cd* Tj = Tn + j * nact;
        for (int a = 0; a < nact; ++a) {
          const cd tj = Tj[a];
#pragma unroll
          for (int e = 0; e < NEX; ++e) z[e] += tj * Rd[e * nact + a];
        }
        // R_e against the occupied columns: T[p_e][n][occ j] Z[j][e]
        const int oj = occ(j);
#pragma unroll
        for (int e = 0; e < NEX; ++e) eX += 2.0 * (Tn[p[e] * nact + oj] * z[e]);
#pragma unroll
        for (int e = 0; e < NEX; ++e)
          if (p[e] == j)
#pragma unroll
            for (int f = 0; f < NEX; ++f) Y[e][f] = z[f];
      }
#pragma unroll
      for (int e = 0; e < NEX; ++e) {
        eX += Y[e][e] * Y[e][e];
        kn += Y[e][e];
#pragma unroll
        for (int f = e + 1; f < NEX; ++f) eX += 2.0 * (Y[f][e] * Y[e][f]);
      }
      // rows whose occupation changed
      for (int i = 0; i < nelec; ++i) {
        const int oi = occ(i), ri = sRef[i];
        const cd tio = Tn[i * nact + oi], tir = Tn[i * nact + ri];
        if (oi != ri) eX += tio * tio - tir * tir;
        for (int j = i + 1; j < nelec; ++j) {
          const int oj = occ(j), rj = sRef[j];
          if (oi != ri || oj != rj)
            eX += 2.0 * (Tn[i * nact + oj] * Tn[j * nact + oi] -
                         Tn[i * nact + rj] * Tn[j * nact + ri]);
        }
        kn += tio;
      }
      KE[((long)d * nw + w) * nchol + n] = kn;
      const cd wd = wgt[(long)d * nw + w];
      if (bad) eX = cmk(__builtin_nan(""), __builtin_nan(""));
      ex_acc += wd * eX;
      ej_acc += wd * (kn * kn - ej0);
    }
  }
  // fixed-order workgroup sums (every thread takes part in both)
  cd sx, sj;
  {
    cd v = ex_acc;
    for (int off = 32; off > 0; off >>= 1) {
      v.re += __shfl_xor(v.re, off, 64);
      v.im += __shfl_xor(v.im, off, 64);
    }
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    sx = cmk(0.0, 0.0);
    for (int i = 0; i < kP2Threads / 64; ++i) sx += red[i];
    __syncthreads();
    v = ej_acc;
    for (int off = 32; off > 0; off >>= 1) {
      v.re += __shfl_xor(v.re, off, 64);
      v.im += __shfl_xor(v.im, off, 64);
    }
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    sj = cmk(0.0, 0.0);
    for (int i = 0; i < kP2Threads / 64; ++i) sj += red[i];
  }
  if (t == 0) {
    part[2 * ((long)w * ngrp + grp)] = sx;
    part[2 * ((long)w * ngrp + grp) + 1] = sj;
  }
}

// EX[w] -= 1/2 sum_g part[w][g].ex, EJ[w] += 1/2 sum_g part[w][g].ej, groups in order
__global__ void k_ph_2body_finish(int nw, int ngrp, const cd* __restrict__ part,
                                  cd* __restrict__ EX, cd* __restrict__ EJ) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  cd sx = cmk(0.0, 0.0), sj = cmk(0.0, 0.0);
  for (int g = 0; g < ngrp; ++g) {
    sx += part[2 * ((long)w * ngrp + g)];
    sj += part[2 * ((long)w * ngrp + g) + 1];
  }
  EX[w] -= 0.5 * sx;
  EJ[w] += 0.5 * sj;
}

// ph_ref_energy_impl's reductions (Real3IndexFactorization.hpp:646-694) over one walker's
// Twina [nel][nchol][nact] and Swia [nel][nact], with the reference at orbitals 0..nel-1 as the
// reference assumes:
//   E[w][0] = E0 + sum_i Swia[i][i]            (only with the one-body term)
//   E[w][1] = -scl/2 sum_{i,j,n} T[i][n][j] T[j][n][i]
//   Kl[w][n] = sum_i T[i][n][i],   E[w][2] = 1/2 sum_n Kl[w][n]^2
// One workgroup per walker, its threads over n; fixed-order workgroup sums.
__global__ __launch_bounds__(256) void k_ph_ref_reduce(int nel, int nchol, int nact, double scl,
                                                       cd E0, int has_h1,
                                                       const cd* __restrict__ Twina,
                                                       const cd* __restrict__ Swia,
                                                       cd* __restrict__ E, cd* __restrict__ Kl) {
  __shared__ cd red[kPhThreads / 64];
  const int w = blockIdx.x;
  const cd* T = Twina + (long)w * nel * nchol * nact;
  cd ex = cmk(0.0, 0.0), ej = cmk(0.0, 0.0), e1 = cmk(0.0, 0.0);
  for (int n = threadIdx.x; n < nchol; n += kPhThreads) {
    cd kl = cmk(0.0, 0.0), x = cmk(0.0, 0.0);
    for (int i = 0; i < nel; ++i) {
      const cd* Ti = T + ((long)i * nchol + n) * nact;
      kl += Ti[i];
      for (int j = 0; j < nel; ++j) x += Ti[j] * T[((long)j * nchol + n) * nact + i];
    }
    Kl[(long)w * nchol + n] = kl;
    ex += x;
    ej += kl * kl;
  }
  if (has_h1)
    for (int i = threadIdx.x; i < nel; i += kPhThreads) e1 += Swia[((long)w * nel + i) * nact + i];
  const cd sx = block_sum(ex, red);
  const cd sj = block_sum(ej, red);
  const cd s1 = block_sum(e1, red);
  if (threadIdx.x == 0) {
    E[3L * w] = has_h1 ? E0 + s1 : cmk(0.0, 0.0);
    E[3L * w + 1] = (-0.5 * scl) * sx;
    E[3L * w + 2] = 0.5 * sj;
  }
}

// The cross-spin coupling of the Cholesky vectors of the excited determinants
// (PHMSD.cpp:470-492): with the couplings C [nout][nin] in CSR form,
//   KEout[d][w][n] = rout[d][w] sum_{u in row d} C[d][u] rin[u][w] KEin[u][w][n]
// (the scaling by the first spin's ratios, csrmm, then the scaling by the second spin's).
// One workgroup per (d, w); threads over n.
__global__ __launch_bounds__(256) void k_ph_couple_ke(int nw, int nchol, const int* __restrict__ cp,
                                                      const int* __restrict__ ci,
                                                      const cd* __restrict__ cv,
                                                      const cd* __restrict__ rin,
                                                      const cd* __restrict__ KEin,
                                                      const cd* __restrict__ rout,
                                                      cd* __restrict__ KEout) {
  const int d = blockIdx.y, w = blockIdx.x;
  const int j0 = cp[d], j1 = cp[d + 1];
  const cd ro = rout ? rout[(long)d * nw + w] : cmk(1.0, 0.0);
  for (int n = threadIdx.x; n < nchol; n += kPhThreads) {
    cd acc = cmk(0.0, 0.0);
    for (int j = j0; j < j1; ++j) {
      const long u = ci[j];
      acc += (cv[j] * rin[u * nw + w]) * KEin[(u * nw + w) * nchol + n];
    }
    KEout[((long)d * nw + w) * nchol + n] = ro * acc;
  }
}

// out[w * ostride] += sum_{d < nd} sum_n a[d][w][n] b[d][w][n]  (PHMSD.cpp:518-519,
// phmsd_helpers.hpp:520-521); one workgroup per walker, fixed-order sums.
__global__ __launch_bounds__(256) void k_ph_ke_dot(int nw, int nchol, int nd,
                                                   const cd* __restrict__ a,
                                                   const cd* __restrict__ b, cd* __restrict__ out,
                                                   int ostride) {
  __shared__ cd red[kPhThreads / 64];
  const int w = blockIdx.x;
  cd acc = cmk(0.0, 0.0);
  for (int d = 0; d < nd; ++d) {
    const long base = ((long)d * nw + w) * nchol;
    for (int n = threadIdx.x; n < nchol; n += kPhThreads) acc += a[base + n] * b[base + n];
  }
  const cd s = block_sum(acc, red);
  if (threadIdx.x == 0) out[(long)w * ostride] += s;
}

// acc[w][c] += Eref[w][c] Ov[w] + Ex[c][w]   (PHMSD.cpp:453-466, 505-526; either term optional)
__global__ void k_ph_accumulate(int nw, const cd* __restrict__ Eref, const cd* __restrict__ Ov,
                                const cd* __restrict__ Ex, cd* __restrict__ acc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * nw) return;
  const int w = e / 3, c = e % 3;
  cd v = acc[e];
  if (Eref) v += Eref[e] * Ov[w];
  if (Ex) v += Ex[(long)c * nw + w];
  acc[e] = v;
}

// E[w][c] = acc[w][c] / Ov[w], Ov -> log Ov + log_ov (PHMSD.cpp:531-543)
__global__ void k_ph_finalize(int nw, const cd* __restrict__ acc, const cd* __restrict__ Ov,
                              const cd* __restrict__ log_ov, cd* __restrict__ E,
                              cd* __restrict__ ov_out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const cd o = Ov[w];
  for (int c = 0; c < 3; ++c) E[3L * w + c] = cdiv(acc[3L * w + c], o);
  ov_out[w] = clog_(o) + log_ov[w];
}

// The multi-determinant G of one spin from its R matrix and the reference G
// (PHMSD.cpp:731-742, 768-771, 775-784): G[w][row0 + a][k] = sum_i R[w][i][a] GA[w][i][k] / Ov[w]
// (the reciprocal of Ov taken once per walker, as the reference's device path does).
__global__ void k_ph_rg(int nel, int nact, int nmo, int ldrows, int row0, const cd* __restrict__ R,
                        const cd* __restrict__ GA, const cd* __restrict__ Ov,
                        cd* __restrict__ G) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y;
  if (e >= (long)nact * nmo) return;
  const int a = (int)(e / nmo), k = (int)(e % nmo);
  const cd* Rw = R + (long)w * nel * nact;
  const cd* Gw = GA + (long)w * nel * nmo;
  cd acc = cmk(0.0, 0.0);
  for (int i = 0; i < nel; ++i) acc += Rw[(long)i * nact + a] * Gw[(long)i * nmo + k];
  G[((long)w * ldrows + row0 + a) * nmo + k] = acc * cdiv(cmk(1.0, 0.0), Ov[w]);
}

// The full G of both spins from the compact one (csrmm 'T' with the orbital rows,
// PHMSD.cpp:744-749, 786-791): Gf[w][s][l][k] = sum_a A_s[a][l] Gc[w][s nact + a][k].
__global__ void k_ph_full_g(int nact, int nmo, const cd* __restrict__ A0,
                            const cd* __restrict__ A1, const cd* __restrict__ Gc,
                            cd* __restrict__ Gf) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y, sp = blockIdx.z;
  if (e >= (long)nmo * nmo) return;
  const int l = (int)(e / nmo), k = (int)(e % nmo);
  const cd* As = sp ? A1 : A0;
  const cd* G = Gc + ((long)w * 2 * nact + (long)sp * nact) * nmo;
  cd acc = cmk(0.0, 0.0);
  for (int a = 0; a < nact; ++a) acc += As[(long)a * nmo + l] * G[(long)a * nmo + k];
  Gf[(((long)w * 2 + sp) * nmo + l) * nmo + k] = acc;
}

// out[w] = log Ov[w] + log_ov[w]
__global__ void k_ph_log_overlap(int nw, const cd* __restrict__ Ov, const cd* __restrict__ log_ov,
                                 cd* __restrict__ out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) out[w] = clog_(Ov[w]) + log_ov[w];
}

__global__ void k_ph_fill(long n, cd v, cd* __restrict__ x) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) x[e] = v;
}

// ------------------------------------------------------------------------------------------------
// PHMSD::energy_alg0 (PHMSD.cpp:122-370), COLLINEAR walkers. Every unique configuration c of a spin
// gets its own compact G (det_ops::MixedDensityMatrixFromConfiguration, density_matrix.hpp:726-775)
// and single-determinant energy over the active rows it occupies (Real3IndexFactorization::energy,
// hpp:176-210 and energy_impl 476-596). With the walker products, formed once per walker by GEMMs,
//   TAB[w][j][a] = sum_k Phi_w[k][j] A[a][k]     Hp[w][j][a] = sum_k Phi_w[k][j] haj[a][k]
//   P[w][j][n][a] = sum_k Phi_w[k][j] Lna[n][a][k]
// the configuration's matrix is T_c[i][j] = TAB[w][j][conf_i] and G_c[a][k] = sum_j Tinv[j][a]
// Phi_w[k][j], so
//   E1 = E0 + sum_{a,j} Tinv[j][a] Hp[w][j][conf_a]
//   Y^n[a][b] = sum_k Lna[n][conf_a][k] G_c[b][k] = sum_j P[w][j][n][conf_a] Tinv[j][b]
//   KE[n] = sum_a Y^n[a][a]    EXX = -1/2 sum_n sum_{a,b} Y^n[a][b] Y^n[b][a]    EJ = 1/2 sum_n KE[n]^2
// and the work per configuration is nel^3 per Cholesky vector rather than nel^2 nmo.

// Phit[w][j][k] = Phi_w[k][j]
__global__ void k_ph0_transpose(int nmo, int nel, const cd* __restrict__ Phi, long ldw,
                                cd* __restrict__ Phit) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y;
  if (e >= (long)nel * nmo) return;
  const int j = (int)(e / nmo), k = (int)(e % nmo);
  Phit[(long)w * nel * nmo + e] = Phi[(long)w * ldw + (long)k * nel + j];
}

__device__ inline void ph0_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline cd ph0_wave_sum(cd v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.re += __shfl_xor(v.re, off, 64);
    v.im += __shfl_xor(v.im, off, 64);
  }
  return v;
}

// One wavefront per (configuration c, walker w). Gauss-Jordan on [T_c | I] with getrf's pivots
// (first max of |re|+|im| in the column among the rows not yet eliminated; the rows at and below
// step k see the same updates as in getrf), log det = sum_k log(+-U_kk + DBL_MIN), the sign
// flipped where rows were swapped, imaginary part wrapped to [-pi, pi]
// (log_determinant_from_getrf). A configuration with an exactly zero pivot gets a zero inverse
// (getri_or_zero) and log det = -inf, i.e. ratio 0.
// Outputs: ld[c][w], Emsd[c][w][0] = E1 and Tinv[c][w][j][b]. Fixed-order sums.
__global__ __launch_bounds__(64) void k_ph0_inverse(
    int nw, int nc, int nel, int nact, const int* __restrict__ conf, const cd* __restrict__ TAB,
    const cd* __restrict__ Hp, cd E0, cd* __restrict__ ld, cd* __restrict__ Emsd,
    cd* __restrict__ Tinv) {
  extern __shared__ cd sm[];
  __shared__ int sConf[kPhMaxElec];
  __shared__ int sPiv, sSing;
  const int t = threadIdx.x;
  const int c = blockIdx.x % nc, w = blockIdx.x / nc;
  const int n2 = 2 * nel;
  cd* sM = sm;                       // [nel][2 nel]: [T | I] -> [I | T^-1]
  cd* sRow = sM + nel * n2;          // [2 nel] the scaled pivot row
  cd* sF = sRow + n2;                // [nel] the pivot column before the update
  if (t < nel) sConf[t] = min(max(conf[(long)c * nel + t], 0), nact - 1);
  __syncthreads();
  const cd* TABw = TAB + (long)w * nel * nact;
  for (int e = t; e < nel * n2; e += 64) {
    const int i = e / n2, j = e % n2;
    sM[e] = j < nel ? TABw[(long)j * nact + sConf[i]] : cmk(j - nel == i ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  cd logd = cmk(0.0, 0.0);           // thread 0
  bool sing = false;                 // uniform: read from LDS after a barrier
  for (int k = 0; k < nel; ++k) {
    {
      double v = -1.0;
      int idx = nel;
      if (t >= k && t < nel) {
        const cd x = sM[t * n2 + k];
        v = fabs(x.re) + fabs(x.im);
        idx = t;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off, 64);
        const int i2 = __shfl_xor(idx, off, 64);
        if (v2 > v || (v2 == v && i2 < idx)) {
          v = v2;
          idx = i2;
        }
      }
      if (t == 0) {
        sPiv = idx;
        sSing = (v == 0.0) ? 1 : 0;
      }
    }
    __syncthreads();
    const int p = sPiv;
    if (sSing) {
      sing = true;
      break;
    }
    if (p != k)
      for (int e = t; e < n2; e += 64) {
        const cd x = sM[k * n2 + e];
        sM[k * n2 + e] = sM[p * n2 + e];
        sM[p * n2 + e] = x;
      }
    __syncthreads();
    const cd piv = sM[k * n2 + k];
    if (t == 0) {
      const cd u = p != k ? -piv : piv;
      logd += clog_(cmk(u.re + DBL_MIN, u.im));
    }
    const cd inv = cdiv(cmk(1.0, 0.0), piv);
    for (int e = t; e < n2; e += 64) sRow[e] = sM[k * n2 + e] * inv;
    for (int e = t; e < nel; e += 64) sF[e] = sM[e * n2 + k];
    __syncthreads();
    for (int e = t; e < nel * n2; e += 64) {
      const int r = e / n2, col = e % n2;
      sM[e] = r == k ? sRow[col] : sM[e] - sF[r] * sRow[col];
    }
    __syncthreads();
  }
  if (sing) {
    for (int e = t; e < nel * nel; e += 64) sM[(e / nel) * n2 + nel + e % nel] = cmk(0, 0);
    __syncthreads();
  }
  if (t == 0) {
    const double pi = 3.14159265358979323846;
    if (sing) {
      logd = cmk(-INFINITY, 0.0);
    } else if (isfinite(logd.im)) {
      while (logd.im > pi) logd.im -= 2 * pi;
      while (logd.im < -pi) logd.im += 2 * pi;
    }
    ld[(long)c * nw + w] = logd;
  }
  const cd* Ti = sM + nel;           // Tinv[j][b] = Ti[j * n2 + b]
  // one-body term
  cd e1 = cmk(0.0, 0.0);
  const cd* Hw = Hp + (long)w * nel * nact;
  for (int e = t; e < nel * nel; e += 64) {
    const int a = e / nel, j = e % nel;
    e1 += Ti[j * n2 + a] * Hw[(long)j * nact + sConf[a]];
  }
  cd* To = Tinv + ((long)c * nw + w) * nel * nel;
  for (int e = t; e < nel * nel; e += 64) To[e] = Ti[(e / nel) * n2 + e % nel];
  const cd s1 = ph0_wave_sum(e1);
  if (t == 0) Emsd[((long)c * nw + w) * 4] = s1 + E0;
}

// The Cholesky terms of one (configuration c, walker w), one wavefront: lane (v, a) forms row a of
// Y^n for vector n = n0 + v, y[b] = sum_j P[w][j][n][conf_a] Tinv[j][b], with Tinv read by scalar
// loads (every lane needs the same element) and P from memory (L2: the workgroups in flight,
// configuration fastest, share the walker's P); the rows go to LDS once so that each lane also
// reads column a for sum_b Y[a][b] Y[b][a], and lane (v, 0) sums the diagonal (KE[n]). Emsd[c][w]
// [1..3] = (EXX, same-spin EJ, sum_n KEr[c][w][n] KE[n]); KE[c][w][n] when KE is given. Sums in a
// fixed order.
template <int NELB, bool EXACT>
__global__ __launch_bounds__(64) void k_ph0_chol(int nw, int nc, int nel_, int nact, int nchol,
                                                 int nvc, const int* __restrict__ conf,
                                                 const cd* __restrict__ Tinv,
                                                 const cd* __restrict__ P,
                                                 cd* __restrict__ Emsd, cd* __restrict__ KE,
                                                 const cd* __restrict__ KEr) {
  extern __shared__ cd sm[];
  const int nel = EXACT ? NELB : nel_;   // a compile-time count when EXACT
  const int lane = threadIdx.x;
  const int c = blockIdx.x % nc, w = blockIdx.x / nc;
  cd* sY = sm;                       // [nvc][nel][nel]
  // uniform across the wavefront and read-only: scalar loads, used as scalar FMA operands
  const cd* __restrict__ Tw = Tinv + ((long)c * nw + w) * nel * nel;
  const int v = lane / nel, a = lane % nel;
  const bool act = lane < nvc * nel;
  const int ca = act ? min(max(conf[(long)c * nel + a], 0), nact - 1) : 0;
  ph0_wave_sync();
  const cd* Pw = P + (long)w * nel * nchol * nact;
  cd exx = cmk(0.0, 0.0), ej = cmk(0.0, 0.0), cross = cmk(0.0, 0.0);
  for (int n0 = 0; n0 < nchol; n0 += nvc) {
    const int n = n0 + v;
    const bool on = act && n < nchol;
    cd y[NELB];
#pragma unroll
    for (int b = 0; b < NELB; ++b) y[b] = cmk(0.0, 0.0);
    if (on) {
      const cd* pj = Pw + (long)n * nact + ca;
#pragma unroll
      for (int j = 0; j < NELB; ++j) {
        if (j < nel) {
          const cd p = pj[(long)j * nchol * nact];
          const cd* ti = Tw + j * nel;
#pragma unroll
          for (int b = 0; b < NELB; ++b)
            if (b < nel) y[b] += p * ti[b];
        }
      }
#pragma unroll
      for (int b = 0; b < NELB; ++b)
        if (b < nel) sY[(v * nel + a) * nel + b] = y[b];
    }
    ph0_wave_sync();
    if (on) {
      cd sx = cmk(0.0, 0.0);
#pragma unroll
      for (int b = 0; b < NELB; ++b)
        if (b < nel) sx += y[b] * sY[(v * nel + b) * nel + a];
      exx += sx;
      if (a == 0) {
        cd kv = cmk(0.0, 0.0);
        for (int d = 0; d < nel; ++d) kv += sY[(v * nel + d) * nel + d];
        ej += kv * kv;
        const long o = ((long)c * nw + w) * nchol + n;
        if (KE) KE[o] = kv;
        if (KEr) cross += KEr[o] * kv;
      }
    }
    ph0_wave_sync();                 // every lane has read sY before the next pass writes it
  }
  exx = ph0_wave_sum(exx);
  ej = ph0_wave_sum(ej);
  cross = ph0_wave_sum(cross);
  if (lane == 0) {
    cd* out = Emsd + ((long)c * nw + w) * 4;
    out[1] = -0.5 * exx;
    out[2] = 0.5 * ej;
    out[3] = cross;
  }
}

// ld[c][w] -> ratio exp(ld[c][w] - ld[0][w]) for c >= 1 (PHMSD.cpp:245-252); row 0 is read here and
// replaced afterwards by k_ph0_ratio_ref
__global__ void k_ph0_ratios(int nw, int nc, cd* __restrict__ r) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x + nw;
  if (e >= (long)nc * nw) return;
  r[e] = cexp_(r[e] - r[e % nw]);
}

// log_ov[w] += ld[0][w], row 0 -> 1 (PHMSD.cpp:245-248, 364-369)
__global__ void k_ph0_ratio_ref(int nw, cd* __restrict__ r, cd* __restrict__ log_ov) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  log_ov[w] += r[w];
  r[w] = cmk(1.0, 0.0);
}

// acc[w][k] = sum_u wgt0[u][w] Em0[u][w][k] + sum_d wgt1[d][w] Em1[d][w][k]
//             (+ sum_d r1[d][w] Em1[d][w][3] for k = 2: the opposite-spin Coulomb term)
// (PHMSD.cpp:301-359); one workgroup per walker, threads over configurations, fixed-order sums.
__global__ __launch_bounds__(kPhThreads) void k_ph0_assemble(
    int nw, int nu, int nd, const cd* __restrict__ wgt0, const cd* __restrict__ Em0,
    const cd* __restrict__ wgt1, const cd* __restrict__ Em1, const cd* __restrict__ r1,
    cd* __restrict__ acc) {
  __shared__ cd red[kPhThreads / 64];
  const int w = blockIdx.x, t = threadIdx.x;
  cd v[3] = {cmk(0.0, 0.0), cmk(0.0, 0.0), cmk(0.0, 0.0)};
  for (int u = t; u < nu; u += kPhThreads) {
    const cd g = wgt0[(long)u * nw + w];
    const cd* e = Em0 + ((long)u * nw + w) * 4;
    for (int k = 0; k < 3; ++k) v[k] += g * e[k];
  }
  for (int d = t; d < nd; d += kPhThreads) {
    const cd g = wgt1[(long)d * nw + w];
    const cd* e = Em1 + ((long)d * nw + w) * 4;
    for (int k = 0; k < 3; ++k) v[k] += g * e[k];
    v[2] += r1[(long)d * nw + w] * e[3];
  }
  for (int k = 0; k < 3; ++k) {
    const cd sk = block_sum(v[k], red);
    if (t == 0) acc[3L * w + k] = sk;
  }
}

}  // namespace

long ph_2body_workspace(int nw, int nact, int nelec, int nchol) {
  // doubles of the LDS kernel's group partials; 0 when its Twina slices do not fit in LDS
  if ((long)kP2N * nelec * nact * 16 > kP2Lds) return 0;
  return 4L * nw * ((nchol + kP2N - 1) / kP2N);
}

void launch_ph_2body(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec, int nchol,
                     const int* iexcit, const int* refc, const double* Twina, const double* R,
                     const double* wgt, double* EX, double* EJ, double* KE, double* ws,
                     long ws_doubles) {
  if (nw <= 0) return;
  AFQ_CHECK(nex >= 1 && nex <= kPhMaxEx, AFQ_EUNSUPPORTED, "ph 2-body energy: nex must be 1..8");
  AFQ_CHECK(nact <= kPhMaxAct && nelec <= kPhMaxElec, AFQ_EUNSUPPORTED,
            "ph 2-body energy: nact <= 256 and nelec <= 64");
  const long need = ph_2body_workspace(nw, nact, nelec, nchol);
  if (ndet > 0 && nchol > 0 && need > 0 && ws != nullptr && ws_doubles >= need) {
    const int ngrp = (nchol + kP2N - 1) / kP2N;
    const size_t lds = (size_t)kP2N * nelec * nact * sizeof(cd);
    auto go2 = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(nw, ngrp), dim3(kP2Threads), lds, s, nw, ndet, nact, nelec,
                         nchol, ngrp, iexcit, refc, reinterpret_cast<const cd*>(Twina),
                         reinterpret_cast<const cd*>(R), reinterpret_cast<const cd*>(wgt),
                         reinterpret_cast<cd*>(ws), reinterpret_cast<cd*>(KE));
    };
    switch (nex) {
      case 1: go2(k_ph_2body_lds<1>); break;
      case 2: go2(k_ph_2body_lds<2>); break;
      case 3: go2(k_ph_2body_lds<3>); break;
      case 4: go2(k_ph_2body_lds<4>); break;
      case 5: go2(k_ph_2body_lds<5>); break;
      case 6: go2(k_ph_2body_lds<6>); break;
      case 7: go2(k_ph_2body_lds<7>); break;
      default: go2(k_ph_2body_lds<8>); break;
    }
    AFQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ph_2body_finish, dim3((nw + 255) / 256), dim3(256), 0, s, nw, ngrp,
                       reinterpret_cast<const cd*>(ws), reinterpret_cast<cd*>(EX),
                       reinterpret_cast<cd*>(EJ));
    AFQ_HIP(hipGetLastError());
    return;
  }
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nw), dim3(kPhThreads), 0, s, nw, ndet, nact, nelec, nchol,
                       iexcit, refc, reinterpret_cast<const cd*>(Twina),
                       reinterpret_cast<const cd*>(R), reinterpret_cast<const cd*>(wgt),
                       reinterpret_cast<cd*>(EX), reinterpret_cast<cd*>(EJ),
                       reinterpret_cast<cd*>(KE));
  };
  switch (nex) {
    case 1: go(k_ph_2body<1>); break;
    case 2: go(k_ph_2body<2>); break;
    case 3: go(k_ph_2body<3>); break;
    case 4: go(k_ph_2body<4>); break;
    case 5: go(k_ph_2body<5>); break;
    case 6: go(k_ph_2body<6>); break;
    case 7: go(k_ph_2body<7>); break;
    default: go(k_ph_2body<8>); break;
  }
  AFQ_HIP(hipGetLastError());
}

void launch_ph_1body(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                     const int* iexcit, const int* refc, const double* S, const double* R,
                     const double* wgt, double* E) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph_1body, dim3(nw), dim3(kPhThreads), 0, s, nw, ndet, nex, nact, nelec,
                     iexcit, refc, reinterpret_cast<const cd*>(S), reinterpret_cast<const cd*>(R),
                     reinterpret_cast<const cd*>(wgt), reinterpret_cast<cd*>(E));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_woodbury(hipStream_t s, int nb, int nmo, int nel, int nact, const double* A,
                        const int* ref, const double* Phi, long ldw, double* G, double* ov,
                        double* QQ0, double* work) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)(((long)nel * nmo + 255) / 256)), dim3(256),
                     0, s, nel, nmo, nact, ref, reinterpret_cast<const cd*>(A),
                     reinterpret_cast<cd*>(work));
  AFQ_HIP(hipGetLastError());
  // T = A[ref] Phi, log det accumulated into ov (wrapped), G = T^-T Phi^T
  launch_green(s, nb, nmo, nel, work, Phi, ldw, G, ov, 2, 1.0, cd{0.0, 0.0}, nullptr, 0.0, 0.0,
               nullptr, 0);
  hipLaunchKernelGGL(k_ph_qq0, dim3((unsigned)(((long)nb * nact * nel + 255) / 256)), dim3(256),
                     0, s, nb, nmo, nel, nact, reinterpret_cast<const cd*>(A),
                     reinterpret_cast<const cd*>(G), reinterpret_cast<cd*>(QQ0));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_spin_weights(hipStream_t s, int nw, int nu, const int* cp, const int* ci,
                            const double* cv, const double* rup, const double* rdn, double* wgt,
                            double* ov) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph_spin_weights, dim3(nw), dim3(kPhThreads), 0, s, nw, nu, cp, ci,
                     reinterpret_cast<const cd*>(cv), reinterpret_cast<const cd*>(rup),
                     reinterpret_cast<const cd*>(rdn), reinterpret_cast<cd*>(wgt),
                     reinterpret_cast<cd*>(ov));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_overlaps(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                        const int* iexcit, const double* T, double* ov) {
  if (nw <= 0 || ndet <= 0) return;
  AFQ_CHECK(nex >= 1 && nex <= kPhMaxEx, AFQ_EUNSUPPORTED, "ph overlaps: nex must be 1..8");
  const unsigned blocks = (unsigned)(((long)nw * ndet + 255) / 256);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, nw, ndet, nact, nelec, iexcit,
                       reinterpret_cast<const cd*>(T), reinterpret_cast<cd*>(ov));
  };
  switch (nex) {
    case 1: go(k_ph_overlaps<1>); break;
    case 2: go(k_ph_overlaps<2>); break;
    case 3: go(k_ph_overlaps<3>); break;
    case 4: go(k_ph_overlaps<4>); break;
    case 5: go(k_ph_overlaps<5>); break;
    case 6: go(k_ph_overlaps<6>); break;
    case 7: go(k_ph_overlaps<7>); break;
    default: go(k_ph_overlaps<8>); break;
  }
  AFQ_HIP(hipGetLastError());
}

void launch_ph_compact_R(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                         const int* iexcit, const int* refc, const double* T, double* R) {
  if (nw <= 0 || ndet <= 0) return;
  AFQ_CHECK(nex >= 1 && nex <= kPhMaxEx, AFQ_EUNSUPPORTED, "ph R matrices: nex must be 1..8");
  const unsigned blocks = (unsigned)(((long)nw * ndet + 255) / 256);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, nw, ndet, nact, nelec, iexcit, refc,
                       reinterpret_cast<const cd*>(T), reinterpret_cast<cd*>(R));
  };
  switch (nex) {
    case 1: go(k_ph_compact_R<1>); break;
    case 2: go(k_ph_compact_R<2>); break;
    case 3: go(k_ph_compact_R<3>); break;
    case 4: go(k_ph_compact_R<4>); break;
    case 5: go(k_ph_compact_R<5>); break;
    case 6: go(k_ph_compact_R<6>); break;
    case 7: go(k_ph_compact_R<7>); break;
    default: go(k_ph_compact_R<8>); break;
  }
  AFQ_HIP(hipGetLastError());
}

void launch_ph_accumulate_R(hipStream_t s, int nw, int ndet, int nex, int nact, int nelec,
                            const int* iexcit, const int* refc, const double* Rc,
                            const double* ov, const double* wgt, double* R) {
  if (nw <= 0) return;
  AFQ_CHECK(nex >= 0 && nex <= kPhMaxEx, AFQ_EUNSUPPORTED, "ph R accumulation: nex must be 0..8");
  hipLaunchKernelGGL(k_ph_accumulate_R, dim3(nw), dim3(256), 0, s, nw, ndet, nex, nact, nelec,
                     iexcit, refc, reinterpret_cast<const cd*>(Rc),
                     reinterpret_cast<const cd*>(ov), reinterpret_cast<const cd*>(wgt),
                     reinterpret_cast<cd*>(R));
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq

namespace afq {

static dim3 ph_blocks(long n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_ph_reference_energy(hipStream_t s, int nw, int nel, int nmo, int nchol, int nact,
                                double scl, double E0, const double* G, const double* Lna,
                                const double* haj, double* E, double* Kl, double* Twina,
                                double* Swia) {
  if (nw <= 0) return;
  // Twina[w][i][n][a] = sum_k G[w][i][k] Lna[n][a][k]; Swia[w][i][a] = sum_k G[w][i][k] haj[a][k]
  launch_gemm_nt_z(s, AFQ_FP64, nw * nel, nchol * nact, nmo, 1.0, G, nmo, Lna, nmo, true, Twina,
                   (long)nchol * nact);
  if (haj)
    launch_gemm_nt_z(s, AFQ_FP64, nw * nel, nact, nmo, 1.0, G, nmo, haj, nmo, true, Swia, nact);
  hipLaunchKernelGGL(k_ph_ref_reduce, dim3(nw), dim3(kPhThreads), 0, s, nel, nchol, nact, scl,
                     cmk(E0, 0.0), haj ? 1 : 0, reinterpret_cast<const cd*>(Twina),
                     reinterpret_cast<const cd*>(Swia), reinterpret_cast<cd*>(E),
                     reinterpret_cast<cd*>(Kl));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_couple_ke(hipStream_t s, int nw, int nchol, int nout, const int* cp, const int* ci,
                         const double* cv, const double* rin, const double* KEin,
                         const double* rout, double* KEout) {
  if (nw <= 0 || nout <= 0) return;
  hipLaunchKernelGGL(k_ph_couple_ke, dim3(nw, nout), dim3(kPhThreads), 0, s, nw, nchol, cp, ci,
                     reinterpret_cast<const cd*>(cv), reinterpret_cast<const cd*>(rin),
                     reinterpret_cast<const cd*>(KEin), reinterpret_cast<const cd*>(rout),
                     reinterpret_cast<cd*>(KEout));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_ke_dot(hipStream_t s, int nw, int nchol, int nd, const double* a, const double* b,
                      double* out, int ostride) {
  if (nw <= 0 || nd <= 0) return;
  hipLaunchKernelGGL(k_ph_ke_dot, dim3(nw), dim3(kPhThreads), 0, s, nw, nchol, nd,
                     reinterpret_cast<const cd*>(a), reinterpret_cast<const cd*>(b),
                     reinterpret_cast<cd*>(out), ostride);
  AFQ_HIP(hipGetLastError());
}

void launch_ph_accumulate(hipStream_t s, int nw, const double* Eref, const double* Ov,
                          const double* Ex, double* acc) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph_accumulate, ph_blocks(3L * nw), dim3(256), 0, s, nw,
                     reinterpret_cast<const cd*>(Eref), reinterpret_cast<const cd*>(Ov),
                     reinterpret_cast<const cd*>(Ex), reinterpret_cast<cd*>(acc));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_finalize(hipStream_t s, int nw, const double* acc, const double* Ov,
                        const double* log_ov, double* E, double* ov_out) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph_finalize, ph_blocks(nw), dim3(256), 0, s, nw,
                     reinterpret_cast<const cd*>(acc), reinterpret_cast<const cd*>(Ov),
                     reinterpret_cast<const cd*>(log_ov), reinterpret_cast<cd*>(E),
                     reinterpret_cast<cd*>(ov_out));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_fill(hipStream_t s, long n, double re, double im, double* x) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_ph_fill, ph_blocks(n), dim3(256), 0, s, n, cmk(re, im),
                     reinterpret_cast<cd*>(x));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_rg(hipStream_t s, int nw, int nel, int nact, int nmo, int ldrows, int row0,
                  const double* R, const double* GA, const double* Ov, double* G) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph_rg, dim3((unsigned)(((long)nact * nmo + 255) / 256), nw), dim3(256), 0,
                     s, nel, nact, nmo, ldrows, row0, reinterpret_cast<const cd*>(R),
                     reinterpret_cast<const cd*>(GA), reinterpret_cast<const cd*>(Ov),
                     reinterpret_cast<cd*>(G));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_full_g(hipStream_t s, int nw, int nact, int nmo, const double* A0,
                      const double* A1, const double* Gc, double* Gf) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph_full_g, dim3((unsigned)(((long)nmo * nmo + 255) / 256), nw, 2),
                     dim3(256), 0, s, nact, nmo, reinterpret_cast<const cd*>(A0),
                     reinterpret_cast<const cd*>(A1), reinterpret_cast<const cd*>(Gc),
                     reinterpret_cast<cd*>(Gf));
  AFQ_HIP(hipGetLastError());
}

void launch_ph_log_overlap(hipStream_t s, int nw, const double* Ov, const double* log_ov,
                           double* out) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph_log_overlap, ph_blocks(nw), dim3(256), 0, s, nw,
                     reinterpret_cast<const cd*>(Ov), reinterpret_cast<const cd*>(log_ov),
                     reinterpret_cast<cd*>(out));
  AFQ_HIP(hipGetLastError());
}

// energy_alg0 (PHMSD.cpp:122-370): the walker products of one spin. Phit, TAB, Hp and P as in
// k_ph0_config; Phi_w = Phi + w ldw (c128 [nmo][nel]).
void launch_ph0_products(hipStream_t s, int nw, int nmo, int nel, int nact, int nchol,
                         const double* Phi, long ldw, const double* A, const double* haj,
                         const double* Lna, double* Phit, double* TAB, double* Hp, double* P) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph0_transpose, dim3((unsigned)(((long)nel * nmo + 255) / 256), nw),
                     dim3(256), 0, s, nmo, nel, reinterpret_cast<const cd*>(Phi), ldw,
                     reinterpret_cast<cd*>(Phit));
  AFQ_HIP(hipGetLastError());
  launch_gemm_nt_z(s, AFQ_FP64, nw * nel, nact, nmo, 1.0, Phit, nmo, A, nmo, true, TAB, nact);
  launch_gemm_nt_z(s, AFQ_FP64, nw * nel, nact, nmo, 1.0, Phit, nmo, haj, nmo, true, Hp, nact);
  launch_gemm_nt_z(s, AFQ_FP64, nw * nel, nchol * nact, nmo, 1.0, Phit, nmo, Lna, nmo, true, P,
                   (long)nchol * nact);
}

void launch_ph0_config(hipStream_t s, int nw, int nc, int nel, int nact, int nchol,
                       const int* conf, const double* TAB, const double* Hp, const double* P,
                       double E0, double* ld, double* Emsd, double* KE, const double* KEr,
                       double* Tinv) {
  if (nw <= 0 || nc <= 0) return;
  AFQ_CHECK(nel >= 1 && nel <= 32, AFQ_EUNSUPPORTED,
            "PHMSD reference energy: more than 32 electrons of one spin");
  const long blocks = (long)nc * nw;
  const size_t lds = sizeof(cd) * ((size_t)nel * 2 * nel + 3 * nel);
  hipLaunchKernelGGL(k_ph0_inverse, dim3((unsigned)blocks), dim3(64), lds, s, nw, nc, nel,
                     nact, conf, reinterpret_cast<const cd*>(TAB),
                     reinterpret_cast<const cd*>(Hp), cmk(E0, 0.0), reinterpret_cast<cd*>(ld),
                     reinterpret_cast<cd*>(Emsd), reinterpret_cast<cd*>(Tinv));
  AFQ_HIP(hipGetLastError());
  const int nvc = 64 / nel;
  const size_t lds2 = sizeof(cd) * (size_t)nvc * nel * nel;
  auto args = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64), lds2, s, nw, nc, nel, nact, nchol,
                       nvc, conf, reinterpret_cast<const cd*>(Tinv),
                       reinterpret_cast<const cd*>(P), reinterpret_cast<cd*>(Emsd),
                       reinterpret_cast<cd*>(KE), reinterpret_cast<const cd*>(KEr));
  };
  if (nel == 10) args(k_ph0_chol<10, true>);   // a compile-time count for the common case
  else if (nel <= 8) args(k_ph0_chol<8, false>);
  else if (nel <= 12) args(k_ph0_chol<12, false>);
  else if (nel <= 16) args(k_ph0_chol<16, false>);
  else args(k_ph0_chol<32, false>);
  AFQ_HIP(hipGetLastError());
}

void launch_ph0_ratios(hipStream_t s, int nw, int nc, double* r, double* log_ov) {
  if (nw <= 0) return;
  if (nc > 1) {
    hipLaunchKernelGGL(k_ph0_ratios, ph_blocks((long)(nc - 1) * nw), dim3(256), 0, s, nw, nc,
                       reinterpret_cast<cd*>(r));
    AFQ_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_ph0_ratio_ref, ph_blocks(nw), dim3(256), 0, s, nw,
                     reinterpret_cast<cd*>(r), reinterpret_cast<cd*>(log_ov));
  AFQ_HIP(hipGetLastError());
}

void launch_ph0_assemble(hipStream_t s, int nw, int nu, int nd, const double* wgt0,
                         const double* Em0, const double* wgt1, const double* Em1,
                         const double* r1, double* acc) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(k_ph0_assemble, dim3(nw), dim3(kPhThreads), 0, s, nw, nu, nd,
                     reinterpret_cast<const cd*>(wgt0), reinterpret_cast<const cd*>(Em0),
                     reinterpret_cast<const cd*>(wgt1), reinterpret_cast<const cd*>(Em1),
                     reinterpret_cast<const cd*>(r1), reinterpret_cast<cd*>(acc));
  AFQ_HIP(hipGetLastError());
}

}  // namespace afq
