"""ctypes binding of the FP64 oracle (liborc.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and only
as the checker. The product library never loads it.

Every function restates a reference HOST_MEMORY routine; the C++ source (afq_oracle.cpp) cites
the reference file:line next to each one.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")

P_WEIGHT, P_PHASE, P_PSEUDO_ELOC, P_E1, P_EXX, P_EJ, P_OVLP = 0, 1, 2, 3, 4, 5, 6
P_LOGSCL_UP, P_LOGSCL_DN, P_IS_UNITARY, P_THETA, P_COUNT = 7, 8, 9, 10, 11


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


_vp = C.c_void_p
_i = C.c_int
_l = C.c_long
_d = C.c_double


def _declare(L):
    sig = {
        "orc_getrf": (_i, [_i, _vp, _vp]),
        "orc_log_det_from_getrf": (None, [_i, _vp, _vp, _vp]),
        "orc_getri_or_zero": (_i, [_i, _vp, _vp]),
        "orc_inverse_logdet": (None, [_i, _i, _vp, _vp, _vp]),
        "orc_log_overlap": (None, [_i, _i, _i, _vp, _vp, _l, _vp]),
        "orc_mixed_density_matrix": (None, [_i, _i, _i, _vp, _vp, _l, _vp, _vp, _i]),
        "orc_half_rotate": (None, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
        "orc_exchange_v0": (None, [_i, _i, _vp, _vp]),
        "orc_vbias": (None, [_i, _i, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp]),
        "orc_vhs": (None, [_i, _i, _i, _d, _vp, _vp, _vp]),
        "orc_energy": (None, [_i, _i, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_exp_hermitian": (None, [_i, _vp, _vp]),
        "orc_one_body_propagator": (None, [_i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_propagate": (None, [_i, _i, _i, _vp, _l, _vp, _vp, _i]),
        "orc_orthogonalize": (None, [_i, _i, _i, _vp, _l, _vp]),
        "orc_construct_X": (None, [_i, _i, _d, _i, _i, _vp, _vp, _vp, _vp]),
        "orc_meanfield_factor": (None, [_i, _i, _vp, _vp, _vp]),
        "orc_walker_update": (None, [_i, _vp, _l, _l, _i, _i, _i, _d, _d, _d, _vp, _vp, _vp,
                                     _vp, _d, _d, _vp, _vp, _vp]),
        "orc_bound_weights": (None, [_i, _vp, _l, _l, _d, _d, _d, _vp]),
        "orc_comb": (_i, [_i, _vp, _d, _i, _vp]),
        "orc_population_control": (_i, [_i, _vp, _l, _l, _d, _vp]),
        "orc_ph_excited_2body": (None, [_i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _vp]),
        "orc_ph_excited_1body": (None, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_ph_overlaps": (None, [_i, _i, _i, _i, _i, _vp, _vp, _vp]),
        "orc_ph_spin_weights": (None, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_ph_woodbury": (None, [_i, _i, _i, _i, _vp, _vp, _vp, _l, _vp, _vp, _vp]),
        "orc_ph_compact_R": (None, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
        "orc_ph_calculate_R": (None, [_i, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp]),
        "orc_system_create": (_vp, [_i, _i, _i, _vp, _vp, _vp, _d, _d]),
        "orc_system_destroy": (None, [_vp]),
        "orc_system_set_params": (None, [_vp, _i, _d, _d, _d, _d, _d, _i, _i, _i]),
        "orc_system_tensors": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_system_mixed_density_matrix": (None, [_vp, _i, _vp, _l, _vp, _vp]),
        "orc_system_log_overlap": (None, [_vp, _i, _vp, _l, _vp]),
        "orc_system_energy": (None, [_vp, _i, _vp, _l, _vp, _vp]),
        "orc_system_seed_properties": (None, [_vp, _i, _vp, _l]),
        "orc_system_step": (None, [_vp, _i, _vp, _l, _vp, _d, _i, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _vp]),
        "orc_system_orthogonalize": (None, [_vp, _i, _vp, _l]),
        "orc_system_stats": (None, [_vp, _vp, _vp]),
        "orc_num_threads": (_i, []),
        "orc_set_num_threads": (None, [_i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    # the returned pointer object holds a reference to the array, so a temporary made by _c()
    # inside the call's argument list lives until the call returns
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt=np.complex128):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------------ det_ops level
def getrf(A):
    A = _c(A).copy()
    n = A.shape[0]
    ipiv = np.zeros(n, np.int32)
    info = lib().orc_getrf(n, _p(A), _p(ipiv))
    return A, ipiv, info


def log_det_from_getrf(LU, ipiv, res=0j):
    LU = _c(LU)
    r = np.array([res], np.complex128)
    lib().orc_log_det_from_getrf(LU.shape[0], _p(LU), _p(np.ascontiguousarray(ipiv, np.int32)),
                                 _p(r))
    return r[0]


def getri_or_zero(LU, ipiv):
    LU = _c(LU).copy()
    info = lib().orc_getri_or_zero(LU.shape[0], _p(LU), _p(np.ascontiguousarray(ipiv, np.int32)))
    return LU, info


def inverse_logdet(A):
    A = _c(A)
    nb, n, _ = A.shape
    inv = np.zeros_like(A)
    ld = np.zeros(nb, np.complex128)
    lib().orc_inverse_logdet(nb, n, _p(A), _p(inv), _p(ld))
    return inv, ld


def log_overlap(PsiT, Phi):
    """det_ops::Log_Overlap: PsiT [N,M], Phi [nb,M,N] -> ov [nb] (accumulated from 0)."""
    PsiT, Phi = _c(PsiT), _c(Phi)
    nb, M, N = Phi.shape
    ov = np.zeros(nb, np.complex128)
    lib().orc_log_overlap(nb, M, N, _p(PsiT), _p(Phi), M * N, _p(ov))
    return ov


def mixed_density_matrix(PsiT, Phi, compact=True):
    PsiT, Phi = _c(PsiT), _c(Phi)
    nb, M, N = Phi.shape
    G = np.zeros((nb, N, M) if compact else (nb, M, M), np.complex128)
    ov = np.zeros(nb, np.complex128)
    lib().orc_mixed_density_matrix(nb, M, N, _p(PsiT), _p(Phi), M * N, _p(G), _p(ov),
                                   int(compact))
    return G, ov


def half_rotate(h, Likn, PsiT):
    h, Likn, PsiT = _c(h, np.float64), _c(Likn, np.float64), _c(PsiT)
    M = h.shape[0]
    K = Likn.shape[2]
    N = PsiT.shape[0]
    haj = np.zeros((N, M), np.complex128)
    Lnak = np.zeros((K, N, M), np.complex128)
    lib().orc_half_rotate(M, N, K, _p(h), _p(Likn), _p(PsiT), _p(haj), _p(Lnak))
    return haj, Lnak


def exchange_v0(Likn):
    Likn = _c(Likn, np.float64)
    M, _, K = Likn.shape
    v0 = np.zeros((M, M))
    lib().orc_exchange_v0(M, K, _p(Likn), _p(v0))
    return v0


def vbias(G, Lnak, dt, walker_type=1, nup=None, ndn=0, Lnak_dn=None):
    G, Lnak = _c(G), _c(Lnak)
    W = G.shape[0]
    K, n_up, M = Lnak.shape
    nup = n_up if nup is None else nup
    v = np.zeros((W, K), np.complex128)
    lib().orc_vbias(W, M, nup, ndn, K, walker_type, dt, _p(Lnak),
                    _p(None if Lnak_dn is None else _c(Lnak_dn)), _p(G), _p(v))
    return v


def vhs(X, Likn, dt):
    X, Likn = _c(X), _c(Likn, np.float64)
    W, K = X.shape
    M = Likn.shape[0]
    V = np.zeros((W, M, M), np.complex128)
    lib().orc_vhs(W, M, K, dt, _p(Likn), _p(X), _p(V))
    return V


def energy(G, haj, Lnak, E0=0.0, walker_type=1, ndn=0, haj_dn=None, Lnak_dn=None):
    G, haj, Lnak = _c(G), _c(haj), _c(Lnak)
    W = G.shape[0]
    K, nup, M = Lnak.shape
    E = np.zeros((W, 3), np.complex128)
    lib().orc_energy(W, M, nup, ndn, K, walker_type, E0, _p(haj),
                     _p(None if haj_dn is None else _c(haj_dn)), _p(Lnak),
                     _p(None if Lnak_dn is None else _c(Lnak_dn)), _p(G), _p(E))
    return E


def exp_hermitian(A):
    A = _c(A)
    out = np.zeros_like(A)
    lib().orc_exp_hermitian(A.shape[0], _p(A), _p(out))
    return out


def propagate(Phi, P1, V, order=6):
    Phi = _c(Phi).copy()
    nb, M, N = Phi.shape
    lib().orc_propagate(nb, M, N, _p(Phi), M * N, _p(_c(P1)), _p(_c(V)), order)
    return Phi


def orthogonalize(Phi):
    Phi = _c(Phi).copy()
    nb, M, N = Phi.shape
    ld = np.zeros(nb, np.complex128)
    lib().orc_orthogonalize(nb, M, N, _p(Phi), M * N, _p(ld))
    return Phi, ld


def back_propagate(fields, pos, nbpsteps, nstab, Refs, P1, Likn, dt, order=6):
    """AFQMCBasePropagator::BackPropagate (AFQMCBasePropagator.cpp:403-532), dense vHS, one Slater
    matrix per reference: fields [W][nfields][K] (the ring), pos the history position (the slot
    the next step would fill), Refs [W][nrefs][M][N]. Newest field set first; each reference
    propagated by Propagate<'H'> (propagate.hpp:241-298 with apply_expM<'H'>, 41-70):
    S <- P1^H S, T = S, T <- (-i/n) V^H T, S += T (n = 1..order), S <- P1^H S; QR
    (orthogonalize) after the steps with ni % nstab == 0 and after the last one, log det R added
    to logdetR. Returns (Refs, logdetR [W][nrefs])."""
    fields = np.asarray(fields, np.complex128)
    W, nfields, _ = fields.shape
    Refs = np.array(Refs, np.complex128)
    nrefs = Refs.shape[1]
    logdetR = np.zeros((W, nrefs), np.complex128)
    P1H = np.asarray(P1, np.complex128).conj().T
    slot = pos % nfields
    for ni in range(nbpsteps - 1, -1, -1):
        slot = nfields - 1 if slot == 0 else slot - 1
        VH = vhs(np.ascontiguousarray(fields[:, slot, :]), Likn, dt).conj().transpose(0, 2, 1)
        for nr in range(nrefs):
            S = np.einsum("ij,wjk->wik", P1H, Refs[:, nr])
            T = S.copy()
            for n in range(1, order + 1):
                T = (-1j / n) * np.einsum("wij,wjk->wik", VH, T)
                S = S + T
            S = np.einsum("ij,wjk->wik", P1H, S)
            if ni == 0 or ni % nstab == 0:
                S, ld = orthogonalize(S)
                logdetR[:, nr] += ld
            Refs[:, nr] = S
    return Refs, logdetR


def construct_X(vbias_, xi, vMF, vbias_bound=50.0, project=False, free_projection=False):
    X = _c(vbias_).copy()
    W, K = X.shape
    HW = np.zeros(W, np.complex128)
    lib().orc_construct_X(W, K, vbias_bound, int(project), int(free_projection), _p(_c(vMF)),
                          _p(_c(xi, np.float64)), _p(X), _p(HW))
    return X, HW


def meanfield_factor(X, vMF):
    X = _c(X)
    W, K = X.shape
    mf = np.zeros(W, np.complex128)
    lib().orc_meanfield_factor(W, K, _p(X), _p(_c(vMF)), _p(mf))
    return mf


def one_body_propagator(h, Likn, vMF, dt):
    """generateP1 (AFQMCBasePropagator.cpp:85-157): (P1, H1) for the mean field vMF."""
    h, L = _c(h, np.float64), _c(Likn, np.float64)
    M, K = h.shape[0], L.shape[2]
    v0 = np.ascontiguousarray(exchange_v0(L), np.float64)
    P1 = np.zeros((M, M), np.complex128)
    H1 = np.zeros((M, M), np.complex128)
    lib().orc_one_body_propagator(M, K, dt, _p(h), _p(v0), _p(L), _p(_c(vMF)), _p(P1), _p(H1))
    return P1, H1


def walker_update(buf, poff, hybrid, dt, Eshift, E0, new_ov, new_E, mf, HW, free_projection=False,
                  use_cp=False, lower_scale=1.0, upper_scale=10.0, weight_floor=100.0,
                  weight_fraction=0.1):
    """walker_update + bound_local_energy (WalkerSetUpdate.hpp:46-191), then the weight bound
    of the step (target population = W), in place on buf [W][ldw]. Returns the weight factors
    and the weights after the update (before the bound) that back-propagation records in its
    history rings (WalkerSetUpdate.hpp:184-190)."""
    W, ldw = buf.shape
    stats = np.zeros(3, np.int64)
    wfac = np.zeros(W, np.complex128)
    wnew = np.zeros(W, np.complex128)
    lib().orc_walker_update(W, _p(buf), ldw, poff, int(hybrid), int(free_projection),
                            int(use_cp), dt, Eshift, E0, _p(_c(new_ov)), _p(_c(new_E)),
                            _p(_c(mf)), _p(_c(HW)), lower_scale, upper_scale, _p(stats),
                            _p(wfac), _p(wnew))
    wst = np.zeros(2, np.int64)
    lib().orc_bound_weights(W, _p(buf), ldw, poff, weight_floor, weight_fraction, float(W),
                            _p(wst))
    return wfac, wnew


def comb(weights, r, num_target=None):
    w = np.ascontiguousarray(weights, np.float64)
    n = len(w) if num_target is None else num_target
    off = np.zeros(len(w), np.int32)
    rc = lib().orc_comb(len(w), _p(w), float(r), n, _p(off))
    return off, rc


def population_control(buf, poff, r):
    """In place on a walker buffer [W, ldw]; returns (offspring, status)."""
    W = buf.shape[0]
    off = np.zeros(W, np.int32)
    rc = lib().orc_population_control(W, _p(buf), buf.shape[1], poff, float(r), _p(off))
    return off, rc


def copy_plan(off):
    """The slot placement of population control on one rank (swap_walkers_async,
    population_control.hpp:143-209) as a source map: src[slot] = the walker copied into slot,
    -1 where the slot keeps its walker. A walker with k offspring keeps its slot; its k-1 copies
    fill the slots of zero-offspring walkers in increasing order. The same map moves every
    per-walker array (WalkerSet::copy_walker copies the back-propagation row too,
    WalkerSetBase.h:413-415)."""
    off = np.asarray(off)
    src = np.full(len(off), -1, np.int64)
    free = iter(np.flatnonzero(off == 0))
    for i, k in enumerate(off):
        for _ in range(1, int(k)):
            src[next(free)] = i
    return src


def apply_copy_plan(rows, src):
    """rows[slot] <- rows[src[slot]] for every slot with a source (in place)."""
    for slot in np.flatnonzero(src >= 0):
        rows[slot] = rows[src[slot]]


def ph_excited_2body(iexcit, refc, Twina, R, wgt):
    """ph_excited_2body_energy_dense_cholesky (phmsd_impl.hpp:29-147): returns the walker sums
    (EX[nw], EJ[nw]) and KE[ndet][nw][nchol]. iexcit [ndet][2][nex] (electron rows, then target
    orbitals), refc [nelec], Twina [nw][nelec][nchol][nact], R [nw][ndet][nex][nact],
    wgt [ndet][nw]."""
    nw, ndet, nex, nact = R.shape
    nelec, nchol = Twina.shape[1], Twina.shape[2]
    iex = np.ascontiguousarray(iexcit, np.int32)
    ref = np.ascontiguousarray(refc, np.int32)
    EX = np.zeros(nw, np.complex128)
    EJ = np.zeros(nw, np.complex128)
    KE = np.zeros((ndet, nw, nchol), np.complex128)
    lib().orc_ph_excited_2body(nw, ndet, nex, nact, nelec, nchol, _p(iex), _p(ref), _p(_c(Twina)),
                               _p(_c(R)), _p(_c(wgt)), _p(EX), _p(EJ), _p(KE))
    return EX, EJ, KE


def ph_excited_1body(iexcit, refc, S, R, wgt):
    """ph_excited_1body_energy (phmsd_impl.hpp:149-183): E[nw] = sum_d wgt sum_e (...)."""
    nw, ndet, nex, nact = R.shape
    nelec = S.shape[1]
    iex = np.ascontiguousarray(iexcit, np.int32)
    ref = np.ascontiguousarray(refc, np.int32)
    E = np.zeros(nw, np.complex128)
    lib().orc_ph_excited_1body(nw, ndet, nex, nact, nelec, _p(iex), _p(ref), _p(_c(S)),
                               _p(_c(R)), _p(_c(wgt)), _p(E))
    return E


def ph_woodbury(A, ref, Phi, ov0=None):
    """MixedDensityMatrixForWoodbury (density_matrix.hpp:651-717), compact: (G [nb][nel][nmo],
    ov [nb] accumulated onto ov0, QQ0 [nb][nact][nel])."""
    nb, nmo, nel = Phi.shape
    nact = A.shape[0]
    G = np.zeros((nb, nel, nmo), np.complex128)
    QQ0 = np.zeros((nb, nact, nel), np.complex128)
    ov = np.zeros(nb, np.complex128) if ov0 is None else np.array(ov0, np.complex128)
    r = np.ascontiguousarray(ref, np.int32)
    lib().orc_ph_woodbury(nb, nmo, nel, nact, _p(_c(A)), _p(r), _p(_c(Phi)), nmo * nel, _p(G),
                          _p(ov), _p(QQ0))
    return G, ov, QQ0


def ph_spin_weights(cp, ci, cv, rup, rdn):
    """PHMSD.cpp:441-450: (wgt [nu][nw], Ov [nw]) from the CSR CI couplings."""
    nu, nw = rup.shape
    cp_, ci_ = np.ascontiguousarray(cp, np.int32), np.ascontiguousarray(ci, np.int32)
    wgt = np.zeros((nu, nw), np.complex128)
    ov = np.zeros(nw, np.complex128)
    lib().orc_ph_spin_weights(nw, nu, _p(cp_), _p(ci_), _p(_c(cv)), _p(_c(rup)), _p(_c(rdn)),
                              _p(wgt), _p(ov))
    return wgt, ov


def ph_overlaps(iexcit, T):
    """calculate_overlaps (phmsd_helpers.hpp:111-195), one shell: ov [ndet][nw]."""
    ndet, two_nex = iexcit.shape
    nw, nact, nelec = T.shape
    iex = np.ascontiguousarray(iexcit, np.int32)
    ov = np.zeros((ndet, nw), np.complex128)
    lib().orc_ph_overlaps(nw, ndet, two_nex // 2, nact, nelec, _p(iex), _p(_c(T)), _p(ov))
    return ov


def ph_compact_R(iexcit, refc, T):
    """get_compact_ph_R_matrices (phmsd_helpers.hpp:317-383): R [nw][ndet][nex][nact]."""
    ndet, two_nex = iexcit.shape
    nex = two_nex // 2
    nw, nact, nelec = T.shape
    iex = np.ascontiguousarray(iexcit, np.int32)
    ref = np.ascontiguousarray(refc, np.int32)
    R = np.zeros((nw, ndet, nex, nact), np.complex128)
    lib().orc_ph_compact_R(nw, ndet, nex, nact, nelec, _p(iex), _p(ref), _p(_c(T)), _p(R))
    return R


def ph_calculate_R(shells, refc, T, wgt):
    """calculate_R (phmsd_helpers.hpp:200-262): shells[k] = iexcit [ndet_k][2 (k + 1)] of rank
    k + 1; wgt [1 + sum ndet_k][nw] (row 0 the reference); R [nw][nelec][nact]."""
    nw, nact, nelec = T.shape
    ndets = np.array([len(x) for x in shells], np.int32)
    iex = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32).reshape(-1)
                                               for x in shells]), np.int32)
    ref = np.ascontiguousarray(refc, np.int32)
    R = np.zeros((nw, nelec, nact), np.complex128)
    lib().orc_ph_calculate_R(nw, len(shells), _p(ndets), nact, nelec, _p(iex), _p(ref),
                             _p(_c(T)), _p(_c(wgt)), _p(R))
    return R


# ------------------------------------------------------------------ random numbers
_ACKLAM_A = (-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
             1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00)
_ACKLAM_B = (-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
             6.680131188771972e+01, -1.328068155288572e+01)
_ACKLAM_C = (-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
             -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00)
_ACKLAM_D = (7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
             3.754408661907416e+00)


def probit(p):
    """Acklam's inverse normal CDF as probit.h:57-135 evaluates it (P. J. Acklam 2002): central
    rational function for 0.02425 <= p <= 0.97575, tail functions outside, 0 outside [0, 1]."""
    p = np.asarray(p, np.float64)
    a, b, c, d = _ACKLAM_A, _ACKLAM_B, _ACKLAM_C, _ACKLAM_D
    out = np.zeros_like(p)
    with np.errstate(divide="ignore", invalid="ignore"):
        lo = (p > 0) & (p < 0.02425)
        q = np.sqrt(-2 * np.log(p[lo]))
        out[lo] = ((((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
                   ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1))
        hi = (p > 0.97575) & (p < 1)
        q = np.sqrt(-2 * np.log(1 - p[hi]))
        out[hi] = -((((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
                    ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1))
        mid = (p >= 0.02425) & (p <= 0.97575)
        q = p[mid] - 0.5
        r = q * q
        out[mid] = ((((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
                    (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1))
    out[p == 0] = -np.inf
    out[p == 1] = np.inf
    return out


def canonical_uniform(raw32):
    """std::uniform_real_distribution<double>(0, 1) over std::mt19937 as libstdc++ computes it
    (generate_canonical<double, 53>): outputs x1, x2 in turn give (x1 + x2 2^32) / 2^64 in double
    arithmetic, the largest double below 1 if that rounds to 1."""
    x = np.asarray(raw32, np.float64).reshape(-1, 2)
    u = (x[:, 0] + x[:, 1] * 4294967296.0) / 18446744073709551616.0
    return np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)


def split_seed(seed, rank, stream):
    """utilities/Random.cpp:7-20: splitmix64 finalizer of seed << 32 | stream << 24 | rank."""
    m = (1 << 64) - 1
    x = ((seed & 0xffffffff) << 32 | (stream & 0xff) << 24 | rank) & m
    x = (x + 0x9e3779b97f4a7c15) & m
    x = ((x ^ (x >> 30)) * 0xbf58476d1ce4e5b9) & m
    x = ((x ^ (x >> 27)) * 0x94d049bb133111eb) & m
    return x ^ (x >> 31)


# ------------------------------------------------------------------ estimators
def estimator_measure(weights, E, ovlp, step, dt):
    """One EnergyEstimator::measure on one rank (EnergyEstimator.h:51-92): the local energies
    [W, 3] contracted with the weights by gemv(transpose(E), w) (no conjugation), the overlap
    dot(exp(ovlp), w), each divided by sum w (MeasurementOutput, Measurements.hpp:96-128);
    ProjectionTime = (step + 1) dt; weight_statistics (Measurements.hpp:130-149)."""
    w = np.asarray(weights, np.complex128)
    E = np.asarray(E, np.complex128)
    den = complex(sum(w[i] for i in range(w.size)))
    avg = [sum(E[i, c] * w[i] for i in range(w.size)) for c in range(3)]
    ov = sum(np.exp(ovlp[i]) * w[i] for i in range(w.size))
    s_abs = sum(abs(w[i]) for i in range(w.size))
    s_abs2 = sum(abs(w[i]) ** 2 for i in range(w.size))
    return {"Energy": (avg[0] + avg[1] + avg[2]) / den, "OnebodyEnergy": avg[0] / den,
            "ExchangeEnergy": avg[1] / den, "CoulombEnergy": avg[2] / den, "Overlap": ov / den,
            "ProjectionTime": (step + 1) * dt,
            "EffectiveNumWalkers": (s_abs * s_abs / s_abs2).real, "Phase": den / s_abs,
            "TotalWeight": den}


def average_pseudo_energy(weights, pseudo, dt):
    """averagePseudoEnergy (average_energy.hpp:53-63): log(sum w exp(dt E) / sum w).real / dt."""
    num = sum(weights[i] * np.exp(dt * pseudo[i]) for i in range(len(weights)))
    x = (num / sum(weights)).real
    # std::log: NaN below 0, -inf at 0
    return float(np.log(x) / dt) if x > 0 else (-np.inf if x == 0 else np.nan)


def average_energy(weights, E1, EXX, EJ):
    """averageEnergy (average_energy.hpp:67-79): sum w (E1 + EXX + EJ) / sum w, real part."""
    num = sum(weights[i] * (E1[i] + EXX[i] + EJ[i]) for i in range(len(weights)))
    return float((num / sum(weights)).real)


def accumulator_bins(samples, binsize):
    """Accumulator::add (Accumulator.hpp:86-101): each bin is the running sum of binsize samples
    divided by binsize; only complete bins are returned (collect_full_bins, 12-33)."""
    bins, cur, fill = [], None, 0
    for x in samples:
        x = np.asarray(x)
        cur = x * 0 + x if cur is None else cur + x
        fill += 1
        if fill >= binsize:
            bins.append(cur / binsize)
            cur, fill = x * 0, 0
    return bins


def eshift_relaxation_factor(equilibration_steps):
    """parameter_defaults.cpp:310-318."""
    return 1.0 - np.exp(-10.0 / equilibration_steps) if equilibration_steps > 1 else 1.0


# ------------------------------------------------------------------ whole system
class System:
    """Closed-shell NOMSD (one determinant) + Real3Index Hamiltonian + base propagator."""

    def __init__(self, h, Likn, PsiT, E0=0.0, dt=0.005):
        self._h = _c(h, np.float64)
        self._L = _c(Likn, np.float64)
        self._P = _c(PsiT)
        self.M = self._h.shape[0]
        self.K = self._L.shape[2]
        self.N = self._P.shape[0]
        self.dt = dt
        self.E0 = E0
        self.ptr = lib().orc_system_create(self.M, self.N, self.K, _p(self._h), _p(self._L),
                                           _p(self._P), E0, dt)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            try:
                _lib.orc_system_destroy(self.ptr)
            except Exception:
                pass
            self.ptr = None

    @property
    def ldw(self):
        return self.M * self.N + P_COUNT

    def set_params(self, taylor_n=6, vbias_bound=50.0, upper_cutoff_scale=10.0,
                   lower_cutoff_scale=1.0, weight_bound_floor=100.0, weight_bound_fraction=0.1,
                   free_projection=False, use_cp_constraint=False, project_force_bias=False):
        lib().orc_system_set_params(self.ptr, taylor_n, vbias_bound, upper_cutoff_scale,
                                    lower_cutoff_scale, weight_bound_floor,
                                    weight_bound_fraction, int(free_projection),
                                    int(use_cp_constraint), int(project_force_bias))

    def tensors(self):
        M, N, K = self.M, self.N, self.K
        out = dict(haj=np.zeros((N, M), np.complex128), Lnak=np.zeros((K, N, M), np.complex128),
                   v0=np.zeros((M, M)), vMF=np.zeros(K, np.complex128),
                   P1=np.zeros((M, M), np.complex128), H1=np.zeros((M, M), np.complex128))
        lib().orc_system_tensors(self.ptr, *[_p(out[k]) for k in
                                             ("haj", "Lnak", "v0", "vMF", "P1", "H1")])
        return out

    def new_walkers(self, W):
        """walker_buffer [W, M*N+11] with every walker at the trial, weight 1, phase 1."""
        buf = np.zeros((W, self.ldw), np.complex128)
        Phi0 = self._P.conj().T
        buf[:, :self.M * self.N] = Phi0.reshape(-1)
        buf[:, self.M * self.N + P_WEIGHT] = 1.0
        buf[:, self.M * self.N + P_PHASE] = 1.0
        buf[:, self.M * self.N + P_IS_UNITARY] = 1.0
        return buf

    def phi(self, buf):
        return buf[:, :self.M * self.N].reshape(-1, self.M, self.N)

    def props(self, buf, p):
        return buf[:, self.M * self.N + p]

    def mixed_density_matrix(self, buf):
        W = buf.shape[0]
        G = np.zeros((W, self.N, self.M), np.complex128)
        ov = np.zeros(W, np.complex128)
        lib().orc_system_mixed_density_matrix(self.ptr, W, _p(buf), buf.shape[1], _p(G), _p(ov))
        return G, ov

    def log_overlap(self, buf):
        W = buf.shape[0]
        ov = np.zeros(W, np.complex128)
        lib().orc_system_log_overlap(self.ptr, W, _p(buf), buf.shape[1], _p(ov))
        return ov

    def energy(self, buf):
        W = buf.shape[0]
        E = np.zeros((W, 3), np.complex128)
        ov = np.zeros(W, np.complex128)
        lib().orc_system_energy(self.ptr, W, _p(buf), buf.shape[1], _p(E), _p(ov))
        return E, ov

    def seed_properties(self, buf):
        lib().orc_system_seed_properties(self.ptr, buf.shape[0], _p(buf), buf.shape[1])

    def step(self, buf, xi, Eshift, hybrid=True, capture=False):
        W = buf.shape[0]
        xi = _c(xi, np.float64)
        out = None
        if capture:
            M, N, K = self.M, self.N, self.K
            out = dict(vbias=np.zeros((W, K), np.complex128), X=np.zeros((W, K), np.complex128),
                       V=np.zeros((W, M, M), np.complex128), ov=np.zeros(W, np.complex128),
                       E=np.zeros((W, 3), np.complex128), wfac=np.zeros(W, np.complex128),
                       weight=np.zeros(W, np.complex128))
            args = [_p(out[k]) for k in ("vbias", "X", "V", "ov", "E", "wfac", "weight")]
        else:
            args = [None] * 7
        lib().orc_system_step(self.ptr, W, _p(buf), buf.shape[1], _p(xi), Eshift, int(hybrid),
                              *args)
        return out

    def orthogonalize(self, buf):
        lib().orc_system_orthogonalize(self.ptr, buf.shape[0], _p(buf), buf.shape[1])

    def stats(self):
        e = np.zeros(3, np.int64)
        w = np.zeros(2, np.int64)
        lib().orc_system_stats(self.ptr, _p(e), _p(w))
        return e, w


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(n)
