"""Python mirror of the walker-step interface over the C ABI (include/afq.h).

Names follow the reference's classes: HamiltonianOperations (Real3IndexFactorization.hpp), the
det_ops free functions (SlaterDeterminantOperations/*.hpp), WalkerSet (WalkerSetBase.h) and
Propagator (AFQMCBasePropagator.h). torch supplies device memory and the current HIP stream;
every computation runs in libsafire_afq.so. There is no CPU fallback: without the library (or a
GPU) the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsafire_afq.so")

CLOSED, COLLINEAR = 1, 2
FP64, FP32, BF16X3, INT8_SPLIT = 0, 1, 2, 3
(P_WEIGHT, P_PHASE, P_PSEUDO_ELOC, P_E1, P_EXX, P_EJ, P_OVLP, P_LOGSCL_UP, P_LOGSCL_DN,
 P_IS_UNITARY, P_THETA, P_COUNT) = range(12)
_ERRORS = {1: "AFQ_EBADSHAPE", 2: "AFQ_EHIP", 3: "AFQ_ENOMEM", 4: "AFQ_EINVAL",
           5: "AFQ_EUNSUPPORTED", 6: "AFQ_ECAPACITY"}


class AfqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


class StepParams(C.Structure):
    """afq_step_params: PropagatorParameters (parameters.hpp:103-126)."""
    _fields_ = [("taylor_n", C.c_int), ("vbias_bound", C.c_double),
                ("upper_cutoff_scale", C.c_double), ("lower_cutoff_scale", C.c_double),
                ("weight_bound_floor", C.c_double), ("weight_bound_fraction", C.c_double),
                ("hybrid", C.c_int), ("free_projection", C.c_int),
                ("use_cp_constraint", C.c_int), ("project_force_bias", C.c_int)]


# Exported symbols of include/afq.h and their argument types.
_vp, _i, _l, _d = C.c_void_p, C.c_int, C.c_long, C.c_double
SIGNATURES = {
    "afq_default_params": (None, [_vp]),
    "afq_last_error": (C.c_char_p, []),
    "afq_version": (_i, []),
    "afq_ham_create": (_i, [_i, _i, _i, _vp, _vp, _vp, _d, _d, _i, _vp]),
    "afq_ham_destroy": (_i, [_vp]),
    "afq_ham_reserve": (_i, [_vp, _i]),
    "afq_ham_set_precision": (_i, [_vp, _i]),
    "afq_ham_set_params": (_i, [_vp, _vp]),
    "afq_ham_get_params": (_i, [_vp, _vp]),
    "afq_ham_tensors": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "afq_ham_stats": (_i, [_vp, _vp, _vp, _vp]),
    "afq_population_control": (_i, [_vp, _i, _vp, _l, _d, _vp]),
    "afq_det_mixed_density_matrix": (_i, [_i, _i, _i, _vp, _vp, _l, _vp, _vp, _vp]),
    "afq_det_propagate": (_i, [_i, _i, _i, _vp, _l, _vp, _vp, _i, _vp]),
    "afq_det_orthogonalize": (_i, [_i, _i, _i, _vp, _l, _vp, _vp]),
    "afq_mixed_density_matrix": (_i, [_vp, _i, _vp, _l, _vp, _vp, _vp]),
    "afq_log_overlap": (_i, [_vp, _i, _vp, _l, _vp, _vp]),
    "afq_vbias": (_i, [_vp, _i, _vp, _vp, _vp]),
    "afq_vhs": (_i, [_vp, _i, _vp, _vp, _vp]),
    "afq_energy": (_i, [_vp, _i, _vp, _vp, _vp]),
    "afq_wfn_energy": (_i, [_vp, _i, _vp, _l, _vp, _vp, _vp]),
    "afq_construct_X": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "afq_propagate": (_i, [_vp, _i, _vp, _l, _vp, _vp]),
    "afq_walker_update": (_i, [_vp, _i, _vp, _l, _vp, _vp, _vp, _vp, _d, _i, _vp]),
    "afq_step": (_i, [_vp, _i, _vp, _l, _vp, _d, _vp]),
    "afq_orthogonalize": (_i, [_vp, _i, _vp, _l, _vp]),
    "afq_set_field_history": (_i, [_vp, _i, _i, _vp, _vp]),
    "afq_history_position": (_i, [_vp, _vp, _vp]),
    "afq_set_weight_history": (_i, [_vp, _vp, _vp, _vp]),
    "afq_back_propagate": (_i, [_vp, _i, _i, _i, _vp, _i, _vp, _vp]),
    "afq_seed_properties": (_i, [_vp, _i, _vp, _l, _vp]),
    "afq_weighted_sums": (_i, [_vp, _i, _vp, _l, _vp, _vp]),
    "afq_estimator_sums": (_i, [_vp, _i, _vp, _l, _vp, _vp, _vp, _vp]),
    "afq_ham_enable_timers": (_i, [_vp, _i]),
    "afq_ham_read_timers": (_i, [_vp, _vp, _vp]),
    "afq_timer_name": (C.c_char_p, [_i]),
    "afq_probe_mfma": (_i, [_i, _vp, _vp]),
    "afq_ph_excited_2body_energy": (_i, [_i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _vp]),
    "afq_ph_excited_2body_energy_noalloc": (_i, [_i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp,
                                                 _vp, _vp, _vp, _vp, _vp]),
    "afq_ph_noalloc_scratch_doubles": (_l, []),
    "afq_ph_excited_1body_energy": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "afq_ph_mixed_density_matrix_woodbury": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _l, _vp, _vp,
                                                  _vp, _vp, _vp]),
    "afq_ph_overlaps": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "afq_ph_spin_weights": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "afq_ph_reference_energy": (_i, [_i, _i, _i, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp]),
    "afq_ph_create": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _d, _i, _vp]),
    "afq_ph_destroy": (None, [_vp]),
    "afq_ph_info": (_i, [_vp, _vp, _vp]),
    "afq_ph_set_energy_algorithm": (_i, [_vp, _i]),
    "afq_ph_energy": (_i, [_vp, _i, _vp, _l, _vp, _vp, _vp]),
    "afq_ph_mixed_density_matrix": (_i, [_vp, _i, _vp, _l, _i, _vp, _vp, _vp]),
    "afq_ph_vbias": (_i, [_vp, _i, _vp, _l, _d, _vp, _vp]),
    "afq_ph_vmf": (_i, [_vp, _d, _vp, _vp]),
    "afq_autotune": (_i, [_vp, _i, C.c_char_p, _vp, _vp]),
    "afq_ph_prepare": (_i, [_vp, _d, _i, _vp]),
    "afq_ph_set_params": (_i, [_vp, _vp]),
    "afq_ph_seed_properties": (_i, [_vp, _i, _vp, _l, _vp]),
    "afq_ph_step": (_i, [_vp, _i, _vp, _l, _vp, _d, _vp]),
    "afq_ph_orthogonalize": (_i, [_vp, _i, _vp, _l, _vp]),
    "afq_ph_population_control": (_i, [_vp, _i, _vp, _l, _d, _vp]),
    "afq_ph_compact_R": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "afq_ph_accumulate_R": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "afq_split_seed": (C.c_ulonglong, [_i, _i, C.c_uint]),
    "afq_rng_create": (_i, [C.c_ulonglong, C.POINTER(_vp)]),
    "afq_rng_destroy": (None, [_vp]),
    "afq_rng_raw": (_i, [_vp, _l, _vp]),
    "afq_rng_uniform": (_i, [_vp, _l, _vp]),
    "afq_probit": (_i, [_l, _vp, _vp, _vp]),
}
NUM_REGIONS = 9

_lib = None


def lib():
    """Loads libsafire_afq.so; raises if it was not built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise AfqError(rc, lib().afq_last_error().decode())


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("safire_amd needs a GPU (torch.cuda.is_available() is False)")
    return torch


def stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def default_params() -> StepParams:
    p = StepParams()
    lib().afq_default_params(C.byref(p))
    return p


def to_device(a, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


class HamiltonianOperations:
    """Real3IndexFactorization + closed-shell NOMSD trial + one-body propagator of one system."""

    def __init__(self, h, Likn, PsiT, E0=0.0, dt=0.005, walker_type=CLOSED):
        _torch()
        self.h = np.ascontiguousarray(h, np.float64)
        self.Likn = np.ascontiguousarray(Likn, np.float64)
        self.PsiT = np.ascontiguousarray(PsiT, np.complex128)
        self.M = self.h.shape[0]
        self.K = self.Likn.shape[2]
        self.N = self.PsiT.shape[0]
        self.dt, self.E0 = dt, E0
        handle = C.c_void_p()
        check(lib().afq_ham_create(self.M, self.N, self.K, self.h.ctypes.data,
                                   self.Likn.ctypes.data, self.PsiT.ctypes.data, E0, dt,
                                   walker_type, C.byref(handle)))
        self.handle = handle
        self.capacity = 0

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().afq_ham_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ldw(self):
        return self.M * self.N + P_COUNT

    def reserve(self, W):
        check(lib().afq_ham_reserve(self.handle, W))
        self.capacity = max(self.capacity, W)

    def autotune(self, W, cache_path=None, stream=None):
        """Picks the fastest GEMM tilings for W walkers (cached in cache_path, a JSON file);
        returns the (vbias, vHS) tilings."""
        self.reserve(W)
        out = (C.c_int * 2)()
        path = None if cache_path is None else os.fsencode(cache_path)
        check(lib().afq_autotune(self.handle, int(W), path, out, stream_ptr(stream)))
        return out[0], out[1]

    def set_precision(self, precision):
        check(lib().afq_ham_set_precision(self.handle, precision))

    def set_params(self, **kw):
        p = self.get_params()
        for k, v in kw.items():
            setattr(p, k, v)
        check(lib().afq_ham_set_params(self.handle, C.byref(p)))

    def get_params(self) -> StepParams:
        p = StepParams()
        check(lib().afq_ham_get_params(self.handle, C.byref(p)))
        return p

    def tensors(self):
        M, N, K = self.M, self.N, self.K
        out = dict(haj=np.zeros((N, M), np.complex128), Lnak=np.zeros((K, N, M), np.complex128),
                   v0=np.zeros((M, M)), vMF=np.zeros(K, np.complex128),
                   P1=np.zeros((M, M), np.complex128), H1=np.zeros((M, M), np.complex128))
        check(lib().afq_ham_tensors(self.handle, *[out[k].ctypes.data for k in
                                                   ("haj", "Lnak", "v0", "vMF", "P1", "H1")]))
        return out

    def enable_timers(self, max_records=4096):
        check(lib().afq_ham_enable_timers(self.handle, max_records))

    def read_timers(self):
        """{region: (total ms, instances)} since the last read (HIP events on the step's stream)."""
        ms = np.zeros(NUM_REGIONS)
        cnt = np.zeros(NUM_REGIONS, np.int64)
        check(lib().afq_ham_read_timers(self.handle, ms.ctypes.data, cnt.ctypes.data))
        return {lib().afq_timer_name(i).decode(): (float(ms[i]), int(cnt[i]))
                for i in range(NUM_REGIONS)}

    def stats(self):
        """(eloc bound total/upper/lower, weight bound total/upper, popctl calls/collapses)."""
        e = np.zeros(3, np.int64)
        w = np.zeros(2, np.int64)
        p = np.zeros(2, np.int64)
        check(lib().afq_ham_stats(self.handle, e.ctypes.data, w.ctypes.data, p.ctypes.data))
        return e, w, p

    # Real3IndexFactorization interface
    def vbias(self, G, v, stream=None):
        check(lib().afq_vbias(self.handle, G.shape[0], ptr(G), ptr(v), stream_ptr(stream)))

    def vHS(self, X, V, stream=None):
        check(lib().afq_vhs(self.handle, X.shape[0], ptr(X), ptr(V), stream_ptr(stream)))

    def energy(self, G, E, stream=None):
        check(lib().afq_energy(self.handle, G.shape[0], ptr(G), ptr(E), stream_ptr(stream)))


class PHMSD:
    """Particle-hole multi-determinant trial over COLLINEAR walkers (PHMSD.hpp; built as
    phmsd_from_params does, Wavefunction.cpp:341-497). occs [ndet][nup + ndn]: sorted occupied
    orbitals (down spin offset by nmo), determinant 0 the reference; ci [ndet]; orbmats
    [nref][nmo][nmo]: trial orbitals as rows; h [nmo][nmo]; Likn [nmo][nmo][nchol]."""

    def __init__(self, occs, ci, orbmats, h, Likn, E0=0.0, walker_block=32):
        _torch()
        self.occs = np.ascontiguousarray(occs, np.int32)
        self.ci = np.ascontiguousarray(ci, np.complex128)
        self.orbmats = np.ascontiguousarray(orbmats, np.complex128)
        if self.orbmats.ndim == 2:
            self.orbmats = self.orbmats[None]
        self.h = np.ascontiguousarray(h, np.float64)
        self.Likn = np.ascontiguousarray(Likn, np.float64)
        self.nmo = self.h.shape[0]
        self.nchol = self.Likn.shape[-1]
        self.nup = int(np.sum(self.occs[0] < self.nmo))
        self.ndn = self.occs.shape[1] - self.nup
        handle = C.c_void_p()
        check(lib().afq_ph_create(self.nmo, self.nup, self.ndn, self.nchol, len(self.ci),
                                  self.occs.ctypes.data, self.ci.ctypes.data,
                                  self.orbmats.shape[0], self.orbmats.ctypes.data,
                                  self.h.ctypes.data, self.Likn.ctypes.data, float(E0),
                                  int(walker_block), C.byref(handle)))
        self.handle = handle
        nu, ns = (C.c_int * 2)(), (C.c_int * 2)()
        check(lib().afq_ph_info(self.handle, nu, ns))
        self.nunique = (nu[0], nu[1])
        self.energy_algorithm = "woodbury"

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().afq_ph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    ENERGY_ALGORITHMS = {"reference": 0, "woodbury": 1}

    def set_energy_algorithm(self, algorithm):
        """The "algorithm" option of PHMSD (PHMSDEnergyAlgorithm, parameters.hpp:32-37):
        "woodbury" (the default, energy_alg1) or "reference" (energy_alg0, PHMSD.cpp:122-370)."""
        if algorithm not in self.ENERGY_ALGORITHMS:
            raise ValueError(f"unknown PHMSD energy algorithm {algorithm!r}")
        check(lib().afq_ph_set_energy_algorithm(self.handle, self.ENERGY_ALGORITHMS[algorithm]))
        self.energy_algorithm = algorithm

    def energy(self, wbuf, E=None, ov=None, stream=None):
        """PHMSD::energy (PHMSD.cpp:89-96: energy_alg1, PHMSD.cpp:377-544, or energy_alg0,
        PHMSD.cpp:122-370, as set_energy_algorithm chose) of the walkers in wbuf [W][ldw]
        (c128, Phi_up [nmo][nup] then Phi_dn [nmo][ndn] per row): (E [W][3], log overlap [W])."""
        torch = _torch()
        W = wbuf.shape[0]
        E = torch.empty((W, 3), dtype=torch.complex128, device=wbuf.device) if E is None else E
        ov = torch.empty(W, dtype=torch.complex128, device=wbuf.device) if ov is None else ov
        check(lib().afq_ph_energy(self.handle, W, ptr(wbuf), wbuf.shape[1], ptr(E), ptr(ov),
                                  stream_ptr(stream)))
        return E, ov

    def MixedDensityMatrix(self, wbuf, full=False, stream=None):
        """PHMSD::MixedDensityMatrix (PHMSD.cpp:549-795): (G, log overlap); compact G
        [W][2 nmo][nmo] or full G [W][2][nmo][nmo]."""
        torch = _torch()
        W, n = wbuf.shape[0], self.nmo
        G = torch.empty((W, 2, n, n) if full else (W, 2 * n, n), dtype=torch.complex128,
                        device=wbuf.device)
        ov = torch.empty(W, dtype=torch.complex128, device=wbuf.device)
        check(lib().afq_ph_mixed_density_matrix(self.handle, W, ptr(wbuf), wbuf.shape[1],
                                                int(bool(full)), ptr(G), ptr(ov),
                                                stream_ptr(stream)))
        return G, ov

    @property
    def ldw(self):
        return self.nmo * (self.nup + self.ndn) + P_COUNT

    def prepare(self, dt, W, stream=None):
        """The propagator for time step dt and the workspace of W walkers."""
        self.dt = float(dt)
        check(lib().afq_ph_prepare(self.handle, self.dt, int(W), stream_ptr(stream)))
        self.params = default_params()

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        check(lib().afq_ph_set_params(self.handle, C.byref(self.params)))

    def new_walkers(self, W):
        """W walkers at the reference determinant of the trial, weight 1 (host, c128)."""
        up = self.orbmats[0][:self.nup].conj().T
        dn = self.orbmats[-1][:self.ndn].conj().T
        buf = np.zeros((W, self.ldw), np.complex128)
        buf[:, :self.nmo * self.nup] = up.reshape(-1)
        buf[:, self.nmo * self.nup:self.nmo * (self.nup + self.ndn)] = dn.reshape(-1)
        off = self.nmo * (self.nup + self.ndn)
        buf[:, off + P_WEIGHT] = 1.0
        buf[:, off + P_PHASE] = 1.0
        buf[:, off + P_IS_UNITARY] = 1.0
        return buf

    def seed(self, wbuf, stream=None):
        check(lib().afq_ph_seed_properties(self.handle, wbuf.shape[0], ptr(wbuf), wbuf.shape[1],
                                           stream_ptr(stream)))

    def step(self, wbuf, xi, Eshift, stream=None):
        check(lib().afq_ph_step(self.handle, wbuf.shape[0], ptr(wbuf), wbuf.shape[1], ptr(xi),
                                float(Eshift), stream_ptr(stream)))

    def Orthogonalize(self, wbuf, stream=None):
        check(lib().afq_ph_orthogonalize(self.handle, wbuf.shape[0], ptr(wbuf), wbuf.shape[1],
                                         stream_ptr(stream)))

    def population_control(self, wbuf, uniform, stream=None):
        check(lib().afq_ph_population_control(self.handle, wbuf.shape[0], ptr(wbuf),
                                              wbuf.shape[1], float(uniform), stream_ptr(stream)))

    def vMF(self, dt, stream=None):
        """PHMSD::vMF (PHMSD.cpp:1329-1565): the mean-field potential [nchol]."""
        torch = _torch()
        v = torch.empty(self.nchol, dtype=torch.complex128, device="cuda")
        check(lib().afq_ph_vmf(self.handle, float(dt), ptr(v), stream_ptr(stream)))
        return v

    def vbias(self, wbuf, dt, v=None, stream=None):
        """PHMSD::vbias (PHMSD.cpp:53-72): v [W][nchol]."""
        torch = _torch()
        W = wbuf.shape[0]
        v = torch.empty((W, self.nchol), dtype=torch.complex128, device=wbuf.device) if v is None else v
        check(lib().afq_ph_vbias(self.handle, W, ptr(wbuf), wbuf.shape[1], float(dt), ptr(v),
                                 stream_ptr(stream)))
        return v


class WalkerSet:
    """walker_buffer c128 [W][M*N + 11] with the reference's layout (WalkerSetBase.cpp:43-124)."""

    def __init__(self, ham: HamiltonianOperations, W: int, buf=None):
        torch = _torch()
        self.ham, self.W = ham, W
        if buf is None:
            Phi0 = ham.PsiT.conj().T.reshape(-1)
            host = np.zeros((W, ham.ldw), np.complex128)
            host[:, :ham.M * ham.N] = Phi0
            host[:, ham.M * ham.N + P_WEIGHT] = 1.0
            host[:, ham.M * ham.N + P_PHASE] = 1.0
            host[:, ham.M * ham.N + P_IS_UNITARY] = 1.0
            buf = host
        self.buf = torch.from_numpy(np.ascontiguousarray(buf, np.complex128)).to("cuda")
        if ham.capacity < W:
            ham.reserve(W)

    @property
    def ldw(self):
        return self.buf.shape[1]

    def SlaterMatrices(self):
        M, N = self.ham.M, self.ham.N
        return self.buf[:, :M * N].reshape(self.W, M, N)

    def set_back_propagation(self, nfields: int, stream=None):
        """The field history of back-propagation (WalkerSet NumBackProp / storeFields): a ring
        Fields [W][nfields][K] c128 that every Propagate fills, position reset to 0;
        nfields = 0 switches it off."""
        torch = _torch()
        self.weight_factors = self.weight_history = None   # unregistered with the old ring
        self.fields = (torch.zeros((self.W, nfields, self.ham.K), dtype=torch.complex128,
                                   device="cuda") if nfields > 0 else None)
        check(lib().afq_set_field_history(self.ham.handle, self.W, int(nfields),
                                          ptr(self.fields) if nfields > 0 else None,
                                          stream_ptr(stream)))

    def set_weight_history(self, stream=None):
        """The weight-factor and weight-history rings [W][3 nfields] c128 of back-propagation
        (getWeightFactors / getWeightHistory), filled by every Propagate and moved by
        population control; needs set_back_propagation first."""
        torch = _torch()
        n = 3 * self.NumBackProp()
        self.weight_factors = torch.empty((self.W, n), dtype=torch.complex128, device="cuda")
        self.weight_history = torch.empty((self.W, n), dtype=torch.complex128, device="cuda")
        check(lib().afq_set_weight_history(self.ham.handle, ptr(self.weight_factors),
                                           ptr(self.weight_history), stream_ptr(stream)))

    def HistoryBufferLength(self):
        return 3 * self.NumBackProp()

    def getWeightFactors(self):
        return getattr(self, "weight_factors", None)

    def getWeightHistory(self):
        return getattr(self, "weight_history", None)

    def NumBackProp(self):
        return 0 if getattr(self, "fields", None) is None else self.fields.shape[1]

    def getFields(self):
        return getattr(self, "fields", None)

    def getHistoryPos(self, stream=None):
        pos = C.c_int()
        check(lib().afq_history_position(self.ham.handle, C.byref(pos), stream_ptr(stream)))
        return pos.value

    def getProperty(self, p):
        return self.buf[:, self.ham.M * self.ham.N + p]

    def host(self):
        return self.buf.cpu().numpy()


class det_ops:
    """SlaterDeterminantOperations free functions on device tensors."""

    @staticmethod
    def MixedDensityMatrix(PsiT, Phi, G, ov, ldw=None, stream=None):
        nb, M, N = Phi.shape
        ld = ldw if ldw is not None else M * N
        check(lib().afq_det_mixed_density_matrix(nb, M, N, ptr(PsiT), ptr(Phi), ld, ptr(G),
                                                 ptr(ov), stream_ptr(stream)))

    @staticmethod
    def Log_Overlap(PsiT, Phi, ov, stream=None):
        nb, M, N = Phi.shape
        check(lib().afq_det_mixed_density_matrix(nb, M, N, ptr(PsiT), ptr(Phi), M * N, None,
                                                 ptr(ov), stream_ptr(stream)))

    @staticmethod
    def Propagate(Phi, P1, V, order=6, stream=None):
        nb, M, N = Phi.shape
        check(lib().afq_det_propagate(nb, M, N, ptr(Phi), M * N, ptr(P1), ptr(V), order,
                                      stream_ptr(stream)))

    @staticmethod
    def orthogonalize(Phi, logdetR=None, stream=None):
        nb, M, N = Phi.shape
        check(lib().afq_det_orthogonalize(nb, M, N, ptr(Phi), M * N, ptr(logdetR),
                                          stream_ptr(stream)))


def capture(fn):
    """Records the HIP work that fn() enqueues on the current stream into a graph and returns
    it (torch.cuda.CUDAGraph; replay() runs it). The walker step, Orthogonalize and population
    control allocate nothing and never synchronise with the host, so a run of steps can be
    captured once its handle has run eagerly (the first calls set kernel attributes)."""
    torch = _torch()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    return g


class Propagator:
    """AFQMCBasePropagator over a WalkerSet: Propagate (one fused step) and Orthogonalize."""

    def __init__(self, ham: HamiltonianOperations):
        self.ham = ham

    def Propagate(self, wset: WalkerSet, xi, Eshift: float, stream=None):
        check(lib().afq_step(self.ham.handle, wset.W, ptr(wset.buf), wset.ldw, ptr(xi),
                             float(Eshift), stream_ptr(stream)))

    def Orthogonalize(self, wset: WalkerSet, stream=None):
        check(lib().afq_orthogonalize(self.ham.handle, wset.W, ptr(wset.buf), wset.ldw,
                                      stream_ptr(stream)))

    def BackPropagate(self, nbpsteps: int, nStabilize: int, wset: WalkerSet, Refs, logdetR,
                      stream=None):
        """AFQMCBasePropagator::BackPropagate (AFQMCBasePropagator.cpp:403-532): Refs
        [W][nrefs][M][N] c128 propagated in place with the adjoint propagator over the last
        nbpsteps stored field sets (newest first); logdetR [W][nrefs] c128 receives the log
        determinants of the orthogonalisations. Needs wset.set_back_propagation first."""
        W, nrefs = Refs.shape[0], Refs.shape[1]
        assert Refs.is_contiguous() and logdetR.is_contiguous()
        assert tuple(logdetR.shape) == (W, nrefs)
        check(lib().afq_back_propagate(self.ham.handle, W, int(nbpsteps), int(nStabilize),
                                       ptr(Refs), nrefs, ptr(logdetR), stream_ptr(stream)))

    def population_control(self, wset: WalkerSet, uniform: float, stream=None):
        check(lib().afq_population_control(self.ham.handle, wset.W, ptr(wset.buf), wset.ldw,
                                           float(uniform), stream_ptr(stream)))

    def seed(self, wset: WalkerSet, stream=None):
        check(lib().afq_seed_properties(self.ham.handle, wset.W, ptr(wset.buf), wset.ldw,
                                        stream_ptr(stream)))

    def weighted_sums(self, wset: WalkerSet, stream=None):
        torch = _torch()
        out = torch.zeros(4, dtype=torch.complex128, device="cuda")
        check(lib().afq_weighted_sums(self.ham.handle, wset.W, ptr(wset.buf), wset.ldw, ptr(out),
                                      stream_ptr(stream)))
        return out

    def estimator_sums(self, wset: WalkerSet, E=None, ov=None, stream=None):
        """c128 [9] on the device: sum E1 w, sum EXX w, sum EJ w, sum w, sum exp(OVLP) w, sum |w|,
        sum |w|^2, sum w exp(dt PSEUDO_ELOC), sum w (E1+EXX+EJ); E [W,3] / ov [W] replace the
        walkers' energy and overlap slots in the first five when given."""
        torch = _torch()
        out = torch.zeros(9, dtype=torch.complex128, device="cuda")
        check(lib().afq_estimator_sums(self.ham.handle, wset.W, ptr(wset.buf), wset.ldw, ptr(E),
                                       ptr(ov), ptr(out), stream_ptr(stream)))
        return out
