"""Writes bh_rhf_closed.npz from the reference's BH inputs and statistical reference,
bh_rhf_closed_snapshot.json from its seeded 20-step snapshot run, and bh_rcasci_rhf_phmsd.npz /
bh_rcasci_rhf_phmsd_snapshot.json from the CASCI PHMSD trial and its seeded snapshot run, read as data with h5dump (no HDF5 module for Python is available here). Run once, in a tree where the
reference exists:

    python tests/golden/make_bh_fixtures.py /root/reference

Inputs (tests/functional/afqmc_inputs/BH): afqmc_H_rhf_closed.h5 -- Hamiltonian/hcore
[1,1,M,1,M], Hamiltonian/DenseFactorized/L [1,1,M,1,M,K], attribute nuclear_energy (read as in
RealDenseHamiltonian.cpp:77-133 and hdf5_helpers.hpp:84-106); afqmc_rhf_nomsd.h5 --
Wavefunction/NOMSD/PsiT_0 in CSR (shape, row_pointers, column_indices, complex values) and
ci_coeffs. Expected outputs (tests/functional/statistical_references/BH/dense_rhf_basis_closed/
rhf_nomsd/closed/results.h5): mean and error of Energy, OnebodyEnergy, ExchangeEnergy,
CoulombEnergy, the run's input (JSON) and num_ranks.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))


def raw(path, dset, dtype, shape):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "x.bin")
        subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path], check=True,
                       stdout=subprocess.DEVNULL)
        a = np.fromfile(out, dtype=dtype)
    return a.reshape(shape)


def text(path, args):
    return subprocess.run([H5DUMP, "-m", "%.17g", *args, path], check=True, capture_output=True,
                          text=True).stdout


def scalar_attr(path, attr):
    s = text(path, ["-a", attr])
    return float(re.search(r"\(0\):\s*([-+0-9.eE]+)", s).group(1))


def complex_array(path, dset):
    """A compound (r, i) dataset, parsed from h5dump's text (-b does not write compounds)."""
    s = text(path, ["-d", dset])
    body = s[s.index("DATA {"):]
    pairs = re.findall(r"\{\s*([-+0-9.eE]+),\s*([-+0-9.eE]+)\s*\}", body)
    return np.array([complex(float(r), float(i)) for r, i in pairs])


def complex_scalar(path, dset):
    s = text(path, ["-d", dset])
    m = re.search(r"\(0\):\s*\{\s*([-+0-9.eE]+),\s*([-+0-9.eE]+)", s)
    return complex(float(m.group(1)), float(m.group(2)))


def main(ref_root):
    base = os.path.join(ref_root, "tests", "functional")
    H = os.path.join(base, "afqmc_inputs", "BH", "afqmc_H_rhf_closed.h5")
    Wf = os.path.join(base, "afqmc_inputs", "BH", "afqmc_rhf_nomsd.h5")
    R = os.path.join(base, "statistical_references", "BH", "dense_rhf_basis_closed", "rhf_nomsd",
                     "closed", "results.h5")
    hdr = text(H, ["-H"])
    M, K = [int(x) for x in re.search(r'DATASET "L".*?\(\s*1,\s*1,\s*(\d+),\s*1,\s*\d+,\s*(\d+)\s*\)',
                                      hdr, re.S).groups()]
    hcore = raw(H, "/Hamiltonian/hcore", "<f8", (M, M))
    L = raw(H, "/Hamiltonian/DenseFactorized/L", "<f8", (M, M, K))
    E0 = scalar_attr(H, "/Hamiltonian/nuclear_energy")
    shape = raw(Wf, "/Wavefunction/NOMSD/PsiT_0/shape", "<i4", (2,))
    rp = raw(Wf, "/Wavefunction/NOMSD/PsiT_0/row_pointers", "<i4", (shape[0] + 1,))
    nnz = int(rp[-1])
    ci = raw(Wf, "/Wavefunction/NOMSD/PsiT_0/column_indices", "<i4", (nnz,))
    vals = complex_array(Wf, "/Wavefunction/NOMSD/PsiT_0/values")
    assert vals.size == nnz
    PsiT = np.zeros(tuple(shape), np.complex128)
    for r in range(shape[0]):
        for j in range(rp[r], rp[r + 1]):
            PsiT[r, ci[j]] = vals[j]
    coeffs = complex_array(Wf, "/Wavefunction/NOMSD/ci_coeffs")
    stats = {}
    for obs in ("Energy", "OnebodyEnergy", "ExchangeEnergy", "CoulombEnergy"):
        for k in ("mean", "error"):
            stats[f"{obs}_{k}"] = complex_scalar(R, f"/measurements/Stage0/{obs}/{k}")
    inp = text(R, ["-d", "/input_file"])
    js = inp[inp.index('"{') + 1:inp.rindex('}"') + 1].replace('\\"', '"')
    run = json.loads(js)
    run["execute"]["wavefunction"]["filename"] = os.path.basename(Wf)
    run["execute"]["hamiltonian"]["filename"] = os.path.basename(H)
    ranks = int(re.search(r"\(0\):\s*(\d+)", text(R, ["-d", "/num_ranks"])).group(1))
    out = os.path.join(HERE, "bh_rhf_closed.npz")
    np.savez_compressed(out, hcore=hcore, Likn=L, E0=E0, PsiT=PsiT,
                        ci=coeffs,
                        stats=json.dumps({k: [v.real, v.imag] for k, v in stats.items()}),
                        run=json.dumps(run), num_ranks=ranks)
    print(f"wrote {out}: M={M} K={K} N={shape[0]} E0={E0} ranks={ranks} "
          f"E={stats['Energy_mean'].real} +- {stats['Energy_error'].real}")


def scalar_value(path, dset):
    """A scalar dataset: a real number, or a complex one stored as an (r, i) compound."""
    s = text(path, ["-d", dset])
    body = s[s.index("DATA {"):]
    m = re.search(r"\(0\):\s*\{\s*([-+0-9.eE]+),\s*([-+0-9.eE]+)", body)
    if m:
        return [float(m.group(1)), float(m.group(2))]
    return [float(re.search(r"\(0\):\s*([-+0-9.eE]+)", body).group(1)), 0.0]


def snapshot(ref_root):
    """The BH closed-shell snapshot: mean and error of every scalar observable of the 20-step,
    50-walker, seed-42 run (tests/functional/snapshot_references/BH/dense_rhf_basis_closed/
    rhf_nomsd/closed/results.h5) and the run's input."""
    R = os.path.join(ref_root, "tests", "functional", "snapshot_references", "BH",
                     "dense_rhf_basis_closed", "rhf_nomsd", "closed", "results.h5")
    obs = {}
    for name in ("Energy", "OnebodyEnergy", "ExchangeEnergy", "CoulombEnergy", "Overlap",
                 "EffectiveNumWalkers", "Phase", "ProjectionTime", "TotalWeight"):
        obs[name] = {k: scalar_value(R, f"/measurements/Stage0/{name}/{k}")
                     for k in ("mean", "error")}
    inp = text(R, ["-d", "/input_file"])
    run = json.loads(inp[inp.index('"{') + 1:inp.rindex('}"') + 1].replace('\\"', '"'))
    for k in ("wavefunction", "hamiltonian"):
        run["execute"][k]["filename"] = os.path.basename(run["execute"][k]["filename"])
    ranks = int(re.search(r"\(0\):\s*(\d+)", text(R, ["-d", "/num_ranks"])).group(1))
    out = os.path.join(HERE, "bh_rhf_closed_snapshot.json")
    with open(out, "w") as f:
        json.dump({"observables": obs, "run": run, "num_ranks": ranks}, f, indent=1)
    print(f"wrote {out}: Energy {obs['Energy']['mean']} +- {obs['Energy']['error']}")


def csr_dense(path, grp):
    shape = raw(path, f"{grp}/shape", "<i4", (2,))
    rp = raw(path, f"{grp}/row_pointers", "<i4", (shape[0] + 1,))
    nnz = int(rp[-1])
    cols = raw(path, f"{grp}/column_indices", "<i4", (nnz,))
    vals = complex_array(path, f"{grp}/values")
    assert vals.size == nnz
    A = np.zeros(tuple(shape), np.complex128)
    for r in range(shape[0]):
        for j in range(rp[r], rp[r + 1]):
            A[r, cols[j]] = vals[j]
    return A


def phmsd_snapshot(ref_root, kind="rhf"):
    """The BH CASCI PHMSD trial (tests/functional/afqmc_inputs/BH/afqmc_casci_rhf_phmsd.h5:
    Wavefunction/PHMSD/PsiT_0 in CSR, occa, occb, ci_coeffs, read as readWfn.cpp:180-215 reads
    them) and the seeded 20-step COLLINEAR snapshot run with it (snapshot_references/BH/
    dense_rhf_basis_closed/rcasci_rhf_phmsd/collinear/results.h5): the scalar observables and
    the run's input. The Hamiltonian is the closed-shell BH one of bh_rhf_closed.npz."""
    base = os.path.join(ref_root, "tests", "functional")
    Wf = os.path.join(base, "afqmc_inputs", "BH", f"afqmc_casci_{kind}_phmsd.h5")
    R = os.path.join(base, "snapshot_references", "BH", "dense_rhf_basis_closed",
                     f"rcasci_{kind}_phmsd", "collinear", "results.h5")
    hdr = text(Wf, ["-H"])
    nd, nup = [int(x) for x in re.search(r'DATASET "occa".*?\(\s*(\d+),\s*(\d+)\s*\)', hdr,
                                         re.S).groups()]
    ndn = int(re.search(r'DATASET "occb".*?\(\s*\d+,\s*(\d+)\s*\)', hdr, re.S).group(1))
    occa = raw(Wf, "/Wavefunction/PHMSD/occa", "<i4", (nd, nup))
    occb = raw(Wf, "/Wavefunction/PHMSD/occb", "<i4", (nd, ndn))
    ci = complex_array(Wf, "/Wavefunction/PHMSD/ci_coeffs")
    if 'GROUP "PsiT_0"' in hdr:
        PsiT0 = csr_dense(Wf, "/Wavefunction/PHMSD/PsiT_0")
    else:
        # occupations of the molecular orbitals only (PHMSDOrbitalType::occ): with the reference
        # determinant on the lowest orbitals of each spin (build_PsiT_MO_phmsd leaves the
        # expansion as it is then) the trial orbitals are the orbital basis itself
        M = int(raw(Wf, "/Wavefunction/PHMSD/occa", "<i4", (nd, nup)).max()) + 1
        assert list(occa[0]) == list(range(nup)) and list(occb[0]) == list(range(ndn))
        PsiT0 = None
    out = os.path.join(HERE, f"bh_rcasci_{kind}_phmsd.npz")
    if PsiT0 is None:
        np.savez_compressed(out, occa=occa, occb=occb, ci=ci)
    else:
        np.savez_compressed(out, PsiT0=PsiT0, occa=occa, occb=occb, ci=ci)
    obs = {}
    for name in ("Energy", "OnebodyEnergy", "ExchangeEnergy", "CoulombEnergy", "Overlap",
                 "EffectiveNumWalkers", "Phase", "ProjectionTime", "TotalWeight"):
        obs[name] = {k: scalar_value(R, f"/measurements/Stage0/{name}/{k}")
                     for k in ("mean", "error")}
    inp = text(R, ["-d", "/input_file"])
    run = json.loads(inp[inp.index('"{') + 1:inp.rindex('}"') + 1].replace('\\"', '"'))
    for k in ("wavefunction", "hamiltonian"):
        run["execute"][k]["filename"] = os.path.basename(run["execute"][k]["filename"])
    ranks = int(re.search(r"\(0\):\s*(\d+)", text(R, ["-d", "/num_ranks"])).group(1))
    snap = os.path.join(HERE, f"bh_rcasci_{kind}_phmsd_snapshot.json")
    with open(snap, "w") as f:
        json.dump({"observables": obs, "run": run, "num_ranks": ranks}, f, indent=1)
    print(f"wrote {out} ({nd} determinants, {nup}+{ndn} electrons) and "
          f"{snap}: Energy {obs['Energy']['mean']} +- {obs['Energy']['error']}")


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    main(root)
    snapshot(root)
    phmsd_snapshot(root, "rhf")
    phmsd_snapshot(root, "uhf")
