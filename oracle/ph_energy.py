"""PHMSD local energy, restated on the CPU -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of the GPU path; the product library never loads
it. It composes the oracle's PHMSD pieces (liborc.so through oracle.py) the way the reference
composes them, and restates in numpy the parts that are plain array algebra:

* build_ph_struct: readWfn.cpp:218-340 (unique excitations per spin and rank, the CI sign of the
  in-place excitation order, configuration indices) and Wavefunction.cpp:371-470 (active-space
  indices, the CI couplings conj(ci) in CSR form, both directions).
* ph_reference_energy: Real3IndexFactorization.hpp:212-233 and ph_ref_energy_impl 598-696.
* energy_alg1: PHMSD.cpp:377-544 (COLLINEAR walkers).
* energy_alg0: PHMSD.cpp:122-370 (COLLINEAR walkers), with Real3IndexFactorization::energy
  (hpp:176-210, 476-596) for each configuration.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


def excitation(ref, confg):
    """get_excitation_number (Excitations.hpp:148-192): (holes as positions in ref, particles as
    orbitals in confg order, sign of sorting ref with the holes replaced in place)."""
    ref = list(ref)
    confg = list(confg)
    holes = [i for i, o in enumerate(ref) if o not in confg]
    parts = [o for o in confg if o not in ref]
    work = list(ref)
    for h, p in zip(holes, parts):
        work[h] = p
    sign = 1
    n = len(work)
    for i in range(n):                       # the reference's exchange count (bubble order)
        for j in range(i + 1, n):
            if work[j] < work[i]:
                sign = -sign
                work[i], work[j] = work[j], work[i]
    return holes, parts, sign


def build_ph_struct(ci, occs, nmo, nup, ndn):
    """The PHMSD structures of a CI expansion. occs [ndet][nup + ndn]: sorted occupied orbitals,
    up in [0, nmo), down in [nmo, 2 nmo); determinant 0 is the reference.

    Returns dict with
      refc:   [refc_up, refc_dn] active indices of the reference orbitals
      shells: [spin][n - 1] = int32 [ndet_n][2 n] unique rank-n excitations (positions, then
              active orbitals), in order of first appearance
      nunique: [nu, nd] (the reference is index 0 of each spin)
      C:      [C_ud (nu x nd), C_du (nd x nu)] as dense complex arrays (conj of the signed CI
              coefficients), and csr: [(cp, ci, cv), (cp, ci, cv)] of the same.
    """
    ci = np.asarray(ci, np.complex128)
    occs = np.asarray(occs)
    ndet = len(ci)
    refs = [list(occs[0, :nup]), list(occs[0, nup:nup + ndn])]
    nel = [nup, ndn]
    unique = [[[] for _ in range(nel[s] + 1)] for s in range(2)]
    configs = []
    for i in range(ndet):
        c = ci[i]
        idx = []
        for s in range(2):
            conf = list(occs[i, :nup]) if s == 0 else list(occs[i, nup:nup + ndn])
            holes, parts, sign = excitation(refs[s], conf)
            c = c * sign
            n = len(holes)
            if n == 0:
                idx.append((0, 0))
                continue
            key = tuple(holes) + tuple(parts)
            if key not in unique[s][n]:
                unique[s][n].append(key)
            idx.append((n, unique[s][n].index(key)))
        configs.append((idx, c))
    # active-space indices: every orbital of the basis is active (Wavefunction.cpp:371-408 maps
    # orbital i and i + nmo to the same index i)
    shells, offsets, nunique = [], [], []
    for s in range(2):
        sh, off, tot = [], [0] * (nel[s] + 1), 1
        for n in range(1, nel[s] + 1):
            off[n] = tot
            lst = [list(k[:n]) + [o - s * nmo for o in k[n:]] for k in unique[s][n]]
            sh.append(np.array(lst, np.int32).reshape(len(lst), 2 * n))
            tot += len(lst)
        shells.append(sh)
        offsets.append(off)
        nunique.append(tot)
    C_ud = np.zeros((nunique[0], nunique[1]), np.complex128)
    entries = [[], []]
    for (iu, idn), c in configs:
        u = 0 if iu[0] == 0 else offsets[0][iu[0]] + iu[1]
        d = 0 if idn[0] == 0 else offsets[1][idn[0]] + idn[1]
        entries[0].append((u, d, np.conj(c)))
        entries[1].append((d, u, np.conj(c)))
        C_ud[u, d] += np.conj(c)
    csr = []
    for s, rows in enumerate(nunique):
        ent = sorted(entries[s], key=lambda e: (e[0], e[1]))
        cp = np.zeros(rows + 1, np.int32)
        for r, _, _ in ent:
            cp[r + 1] += 1
        cp = np.cumsum(cp).astype(np.int32)
        csr.append((cp, np.array([e[1] for e in ent], np.int32),
                    np.array([e[2] for e in ent], np.complex128)))
    refc = [np.array(refs[0], np.int32), np.array([o - nmo for o in refs[1]], np.int32)]
    return {"refc": refc, "shells": shells, "nunique": nunique, "C": [C_ud, C_ud.T.copy()],
            "csr": csr}


def ph_reference_energy(G, Lna, haj, E0=0.0, scl=1.0):
    """ph_reference_energy / ph_ref_energy_impl (Real3IndexFactorization.hpp:212-233, 598-696):
    G [nw][nel][nmo] (the reference determinant's compact G), Lna [nchol][nact][nmo],
    haj [nact][nmo] or None (no one-body term). Returns (E [nw][3], Kl [nw][nchol],
    Twina [nw][nel][nchol][nact], Swia [nw][nel][nact] or None). Assumes refc = 0..nel-1, as the
    reference does (E1 takes the diagonal of Swia, the exchange the first nel columns)."""
    nw, nel, _ = G.shape
    E = np.zeros((nw, 3), np.complex128)
    Swia = None
    if haj is not None:
        E[:, 0] = E0
        Swia = np.einsum("wik,ak->wia", G, haj, optimize=True)
        E[:, 0] += np.einsum("wii->w", Swia[:, :, :nel])
    Twina = np.einsum("wik,nak->wina", G, Lna, optimize=True)
    Tn = Twina[:, :, :, :nel]
    E[:, 1] = -0.5 * scl * np.einsum("winj,wjni->w", Tn, Tn, optimize=True)
    Kl = np.einsum("wini->wn", Tn)
    E[:, 2] = 0.5 * np.einsum("wn,wn->w", Kl, Kl)
    return E, Kl, Twina, Swia


def _shell_overlaps(struct, s, QQ0, nw):
    ratios = np.zeros((struct["nunique"][s], nw), np.complex128)
    ratios[0] = 1.0
    row = 1
    for sh in struct["shells"][s]:
        if len(sh):
            ratios[row:row + len(sh)] = O.ph_overlaps(sh, QQ0)
        row += len(sh)
    return ratios


def _excited(struct, s, wgt, QQ0, Twina, Swia, KE):
    """ph_excited_energies_{first,second}_step (phmsd_helpers.hpp:408-523) without the
    determinant blocks (numerically irrelevant): E planar [3][nw] and KE rows per determinant."""
    nw = QQ0.shape[0]
    E = np.zeros((3, nw), np.complex128)
    refc = struct["refc"][s]
    row = 1
    for sh in struct["shells"][s]:
        nd = len(sh)
        if nd:
            R = O.ph_compact_R(sh, refc, QQ0)
            w = np.ascontiguousarray(wgt[row:row + nd])
            E[0] += O.ph_excited_1body(sh, refc, Swia, R, w)
            EX, EJ, K = O.ph_excited_2body(sh, refc, Twina, R, w)
            E[1] += EX
            E[2] += EJ
            KE[row:row + nd] = K
        row += nd
    return E


def energy_alg1(struct, A, Lna, haj, E0, Phi_up, Phi_dn):
    """PHMSD::energy_alg1 (PHMSD.cpp:377-544) for COLLINEAR walkers. A[s] [nact][nmo] (the
    trial orbitals as rows), Lna[s] [nchol][nact][nmo], haj[s] [nact][nmo], Phi_s [nw][nmo][nel_s].
    Returns (E [nw][3], log overlap [nw])."""
    nw = Phi_up.shape[0]
    nchol = Lna[0].shape[0]
    nu, nd = struct["nunique"]
    refc = struct["refc"]
    G0, log_ov, Q0 = O.ph_woodbury(A[0], refc[0], Phi_up)
    G1, log_ov, Q1 = O.ph_woodbury(A[1], refc[1], Phi_dn, log_ov)
    r_up = _shell_overlaps(struct, 0, Q0, nw)
    r_dn = _shell_overlaps(struct, 1, Q1, nw)
    Eacc = np.zeros((nw, 3), np.complex128)
    # spin up
    wgt, Ov = O.ph_spin_weights(*struct["csr"][0], r_up, r_dn)
    Eref, Kl, Twina, Swia = ph_reference_energy(G0, Lna[0], haj[0], E0)
    Eacc += Eref * Ov[:, None]
    KE = np.zeros((max(nu, nd), nw, nchol), np.complex128)
    KE[0] = Kl
    Eacc += _excited(struct, 0, wgt, Q0, Twina, Swia, KE).T
    Tdn = KE[:nu] * r_up[:, :, None]
    KEr = np.einsum("du,uwn->dwn", struct["C"][1], Tdn, optimize=True) * r_dn[:, :, None]
    # spin down
    wgt1, _ = O.ph_spin_weights(*struct["csr"][1], r_dn, r_up)
    Eref, KEl, Twina, Swia = ph_reference_energy(G1, Lna[1], haj[1], 0.0)
    Eacc += Eref * Ov[:, None]
    Eacc[:, 2] += np.einsum("wn,wn->w", KEr[0], KEl)
    KEd = np.zeros((nd, nw, nchol), np.complex128)
    Ed = _excited(struct, 1, wgt1, Q1, Twina, Swia, KEd)
    Ed[2] += np.einsum("dwn,dwn->w", KEr[1:nd], KEd[1:nd])
    Eacc += Ed.T
    return Eacc / Ov[:, None], np.log(Ov) + log_ov


def _config_energy(G, Lna, haj, E0):
    """Real3IndexFactorization::energy (hpp:176-210) with energy_impl (476-596), COLLINEAR, over
    G [nw][nact][nmo] zero outside the configuration's rows: (E [nw][3], EJn [nw][nchol])."""
    E = np.zeros((G.shape[0], 3), np.complex128)
    E[:, 0] = E0 + np.einsum("wak,ak->w", G, haj, optimize=True)
    T = np.einsum("wbk,nak->wbna", G, Lna, optimize=True)          # Twbna
    E[:, 1] = -0.5 * np.einsum("wanb,wbna->w", T, T, optimize=True)
    EJn = np.einsum("wak,nak->wn", G, Lna, optimize=True)
    E[:, 2] = 0.5 * np.einsum("wn,wn->w", EJn, EJn)
    return E, EJn


def energy_alg0(struct, A, Lna, haj, E0, Phi_up, Phi_dn):
    """PHMSD::energy_alg0 (PHMSD.cpp:122-370) for COLLINEAR walkers, arguments as energy_alg1:
    for each spin and unique configuration (get_configuration), the configuration's compact G and
    log overlap (MixedDensityMatrixFromConfiguration, density_matrix.hpp:726-775: the same getrf,
    log_determinant_from_getrf and getri_or_zero as ph_woodbury over the configuration's rows)
    scattered into the active rows, its energy and Cholesky trace (Real3IndexFactorization::
    energy), the overlap ratio exp(log ov_n - log ov_0); the up spin's traces scaled by their
    ratios and carried to the down configurations through the couplings (csrmm); the CI-weighted
    sums, the opposite-spin Coulomb term, division by the overlap. Returns (E [nw][3], log ov)."""
    Phis = [Phi_up, Phi_dn]
    nw = Phi_up.shape[0]
    nchol = Lna[0].shape[0]
    nact, nmo = A[0].shape
    nex = struct["nunique"]
    maxc = max(nex)
    log_ov = np.zeros((2, nw), np.complex128)
    Ovmsd = np.zeros((2, maxc, nw), np.complex128)
    Emsd = np.zeros((2, maxc, nw, 3), np.complex128)
    KEright = np.zeros((maxc, nw, nchol), np.complex128)
    opSpinEJ = np.zeros(nw, np.complex128)
    for s in range(2):
        for nd, confg in enumerate(configurations(struct, s)):
            Gw, ov, _ = O.ph_woodbury(A[s], confg, Phis[s])
            if nd == 0:
                log_ov[s] = ov
                Ovmsd[s, 0] = 1.0
            else:
                Ovmsd[s, nd] = np.exp(ov - log_ov[s])
            G = np.zeros((nw, nact, nmo), np.complex128)
            G[:, confg, :] = Gw
            Ec, EJn = _config_energy(G, Lna[s], haj[s], E0 if s == 0 else 0.0)
            Emsd[s, nd] = Ec
            if s == 0:
                KEright[nd] = EJn
            else:
                opSpinEJ += np.einsum("wn,wn->w", KEright[nd], EJn) * Ovmsd[1, nd]
        if s == 0:
            KEright[:nex[0]] *= Ovmsd[0, :nex[0], :, None]
            Tdn = KEright[:nex[0]].copy()
            KEright[:nex[1]] = np.einsum("du,uwn->dwn", struct["C"][1], Tdn, optimize=True)
    wgt = struct["C"][0] @ Ovmsd[1, :nex[1]]
    Ov = np.einsum("uw,uw->w", Ovmsd[0, :nex[0]], wgt)
    wgt = wgt * Ovmsd[0, :nex[0]]
    Emsd[0, :nex[0]] *= wgt[:, :, None]
    wgt = (struct["C"][1] @ Ovmsd[0, :nex[0]]) * Ovmsd[1, :nex[1]]
    Emsd[1, :nex[1]] *= wgt[:, :, None]
    E = Emsd.sum(axis=(0, 1))
    E[:, 2] += opSpinEJ
    E /= Ov[:, None]
    return E, np.log(Ov) + log_ov[1] + log_ov[0]


def mixed_density_matrix(struct, A, Phi_up, Phi_dn, compact=True):
    """PHMSD::MixedDensityMatrix (PHMSD.cpp:549-795), COLLINEAR walkers. Returns (G, log overlap):
    compact G [nw][nact_up + nact_dn][nmo] (rows: the active orbitals of each spin,
    G_s = R_s^T G_ref,s / Ov with R_s = calculate_R of the spin's determinant weights), or the full
    G [nw][2][nmo][nmo] = A_s^T G_s (csrmm 'T' with the orbital rows)."""
    refc = struct["refc"]
    nw = Phi_up.shape[0]
    G0, log_ov, Q0 = O.ph_woodbury(A[0], refc[0], Phi_up)
    G1, log_ov, Q1 = O.ph_woodbury(A[1], refc[1], Phi_dn, log_ov)
    r_up = _shell_overlaps(struct, 0, Q0, nw)
    r_dn = _shell_overlaps(struct, 1, Q1, nw)
    wgt0, Ov = O.ph_spin_weights(*struct["csr"][0], r_up, r_dn)
    Ra = O.ph_calculate_R(struct["shells"][0], refc[0], Q0, wgt0)
    wgt1, _ = O.ph_spin_weights(*struct["csr"][1], r_dn, r_up)
    Rb = O.ph_calculate_R(struct["shells"][1], refc[1], Q1, wgt1)
    Gc = np.concatenate([np.einsum("wia,wik->wak", Ra, G0), np.einsum("wia,wik->wak", Rb, G1)],
                        axis=1)
    Gc = Gc * (1.0 / Ov)[:, None, None]
    logov = np.log(Ov) + log_ov
    if compact:
        return Gc, logov
    na = A[0].shape[0]
    Gf = np.stack([np.einsum("al,wak->wlk", A[0], Gc[:, :na]),
                   np.einsum("al,wak->wlk", A[1], Gc[:, na:])], axis=1)
    return Gf, logov


def vbias(struct, A, Lna, dt, Phi_up, Phi_dn):
    """PHMSD::vbias (PHMSD.cpp:53-72) with Real3IndexFactorization::vbias (hpp:328-361) for
    COLLINEAR walkers: v[w][n] = sqrt(dt) sum_s sum_{a,k} Lna_s[n][a][k] G_s[w][a][k] over the
    compact multi-determinant G."""
    Gc, _ = mixed_density_matrix(struct, A, Phi_up, Phi_dn)
    na = A[0].shape[0]
    return np.sqrt(dt) * (np.einsum("wak,nak->wn", Gc[:, :na], Lna[0], optimize=True) +
                          np.einsum("wak,nak->wn", Gc[:, na:], Lna[1], optimize=True))


def configurations(struct, s):
    """get_configuration (Excitations.hpp:486-492): the reference occupation with each unique
    excitation of spin s applied in place (index 0 = the reference)."""
    ref = [int(o) for o in struct["refc"][s]]
    out = [list(ref)]
    for sh in struct["shells"][s]:
        n = sh.shape[1] // 2
        for e in sh:
            c = list(ref)
            for p in range(n):
                c[int(e[p])] = int(e[n + p])
            out.append(c)
    return out


def vmf(struct, A, Lna, dt):
    """PHMSD::vMF (PHMSD.cpp:1329-1565), COLLINEAR, as the reference computes it: overlaps of
    every unique configuration relative to the reference (log det(As^H As)); the diagonal term of
    each configuration (its own compact G, weighted by its overlap times sum_n |coupling|^2 x the
    other spin's overlap); the off-diagonal term of every pair of one spin's configurations that
    share a row of the other spin's couplings and differ by one orbital (weight c1 conj(c2), the
    orbital rows added as the reference adds them, no fermionic sign); division by
    sum |c|^2 ov_up ov_dn; then vbias with sqrt(dt)."""
    nact, nmo = A[0].shape
    confs = [configurations(struct, 0), configurations(struct, 1)]
    ratio = []
    for s in range(2):
        Orbs = A[s].conj().T
        logs = []
        for c in confs[s]:
            As = Orbs[:, c]
            sign, ld = np.linalg.slogdet(As.conj().T @ As)
            logs.append(ld + np.log(sign))
        logs = np.array(logs)
        r = np.exp(logs - logs[0])
        r[0] = 1.0
        ratio.append(r)
    cp0, ci0, cv0 = struct["csr"][0]
    ov = 0.0
    for u in range(len(cp0) - 1):
        for j in range(cp0[u], cp0[u + 1]):
            ov += abs(cv0[j]) ** 2 * ratio[0][u] * ratio[1][ci0[j]]
    G = np.zeros((2, nact, nmo), np.complex128)
    for s in range(2):
        o = 1 - s
        Orbs = A[s].conj().T
        cp, ci, cv = struct["csr"][s]
        for nd, c in enumerate(confs[s]):
            As = Orbs[:, c]
            G2 = np.linalg.inv(As.conj().T @ As).T @ As.T
            w = ratio[s][nd] * sum(ratio[o][ci[j]] * abs(cv[j]) ** 2 for j in range(cp[nd], cp[nd + 1]))
            for i, orb in enumerate(c):
                G[s, orb] += w * G2[i]
        cpo, cio, cvo = struct["csr"][o]
        for nd in range(len(cpo) - 1):
            for n1 in range(cpo[nd], cpo[nd + 1]):
                c1 = sorted(confs[s][cio[n1]])
                for n2 in range(n1 + 1, cpo[nd + 1]):
                    c2 = sorted(confs[s][cio[n2]])
                    holes = [i for i, x in enumerate(c1) if x not in c2]
                    if len(holes) != 1:
                        continue
                    part = [x for x in c2 if x not in c1][0]
                    w = cvo[n1] * np.conj(cvo[n2])
                    o0 = c1[holes[0]]
                    G[s, part] += w * A[s][o0]
                    G[s, o0] += np.conj(w) * A[s][part]
    G /= ov
    return np.sqrt(dt) * (np.einsum("ak,nak->n", G[0], Lna[0]) + np.einsum("ak,nak->n", G[1], Lna[1]))


def ph_step(struct, A, Lna, haj, E0, Likn, vMF, P1, dt, buf, nmo, nup, ndn, xi, Eshift,
            hybrid, taylor_n=6, vbias_bound=50.0):
    """AFQMCBasePropagator::Propagate (AFQMCBasePropagator.cpp:289-392) with the PHMSD trial over
    COLLINEAR walkers: vbias (PHMSD::vbias), assemble_X and the mean-field factor, vHS, both
    spins' Slater matrices propagated with the same P1 and vHS, the new overlap (hybrid) or the
    PHMSD local energy and overlap, walker_update and the weight bound. Returns the new buffer."""
    buf = np.array(buf, np.complex128)
    W = buf.shape[0]
    nu = nmo * nup
    poff = nmo * (nup + ndn)
    Pu = buf[:, :nu].reshape(W, nmo, nup)
    Pd = buf[:, nu:poff].reshape(W, nmo, ndn)
    v = vbias(struct, A, Lna, dt, Pu, Pd)
    X, HW = O.construct_X(v, xi, vMF, vbias_bound)
    mf = O.meanfield_factor(X, vMF)
    V = O.vhs(X, Likn, dt)
    Pu = O.propagate(Pu, P1, V, taylor_n)
    Pd = O.propagate(Pd, P1, V, taylor_n)
    buf[:, :nu] = Pu.reshape(W, -1)
    buf[:, nu:poff] = Pd.reshape(W, -1)
    if hybrid:
        _, new_ov = mixed_density_matrix(struct, A, Pu, Pd)
        new_E = np.zeros((W, 3), np.complex128)
    else:
        new_E, new_ov = energy_alg1(struct, A, Lna, haj, E0, Pu, Pd)
    O.walker_update(buf, poff, hybrid, dt, Eshift, E0, new_ov, new_E, mf, HW)
    return buf
