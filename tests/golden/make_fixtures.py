"""Writes the golden fixtures of tests/golden from the reference's own test
inputs and known answers (run here, where the MDFF tree is mounted; the
fixtures are data and travel with the repository).

  ex0_lj_fcc.npz     examples/ex0/POSFF (256-atom fcc Ar), DL_POLY energy
                     examples/ex0/dl_poly/STATIS:4, LR correction
                     examples/ex0/dl_poly/OUTPUT:124
  ex2_clusters.npz   examples/ex2/clusters/{N} and the Cambridge Cluster
                     Database energies of examples/ex2/REFERENCE
  ex4_efg.json       examples/ex4/TRAJFF.twocharges, TRAJFF.cluster, the
                     Nymand-Linse values (run_ex4.sh:78-79) and the GULP
                     principal EFGs and asymmetries (run_ex4.sh:161-169)
  ex9_pim.npz/json   examples/ex9/config/POSFF and pot_mdff/*.POT (test 11:
                     BMHFTD + damped PIM + Ewald), with the reference's
                     Utot and SCF iteration count from a run of mdff.x on
                     that input
"""
import json
import os
import re

import numpy as np

REF = "/root/reference/examples"
OUT = os.path.dirname(os.path.abspath(__file__))


def ex0():
    lines = open(f"{REF}/ex0/POSFF").read().split("\n")
    L = float(lines[2].split()[0])
    n = int(lines[0])
    frac = np.array([[float(x) for x in l.split()[1:4]] for l in lines[9:9 + n]])
    statis = open(f"{REF}/ex0/dl_poly/STATIS").read().split("\n")[3].split()
    np.savez(f"{OUT}/ex0_lj_fcc.npz", frac=frac, L=L, eps=0.010323576, sigma=3.405, rc=10.0,
             dlpoly_eng_tot=float(statis[0]), dlpoly_lr_energy=-0.954176)


def ex2():
    ref = {}
    for l in open(f"{REF}/ex2/REFERENCE"):
        t = l.split()
        if len(t) == 2 and not l.startswith("#"):
            ref[t[0]] = float(t[1])
    out = {}
    for name in ["13", "38", "55", "75", "98", "100", "147", "38i", "75i"]:
        xyz = np.loadtxt(f"{REF}/ex2/clusters/{name}")
        out[f"pos_{name}"] = xyz
        out[f"e_{name}"] = ref[name]
    np.savez(f"{OUT}/ex2_clusters.npz", **out)


def _traj(path):
    t = open(path).read().split("\n")
    n = int(t[0])
    A = np.array([[float(x) for x in t[k].split()] [:3] for k in (2, 3, 4)]).T
    names = t[6].split()
    counts = [int(x) for x in t[7].split()]
    rows = [l.split() for l in t[9:9 + n]]
    pos = [[float(x) for x in r[1:4]] for r in rows]
    typ = [names.index(r[0]) + 1 for r in rows]
    return {"A": A.tolist(), "pos": pos, "itype": typ, "counts": counts}


def ex4():
    gulp = [[15.2454, 16.4848, -31.7301], [13.1974, 15.5994, -28.7968],
            [20.5185, 23.5380, -44.0565], [12.0418, 14.6251, -26.6669],
            [8.6231, 10.8748, -19.4979], [22.1626, 24.1320, -46.2946],
            [7.7041, 9.9389, -17.6430], [16.7300, 17.0914, -33.8214],
            [-0.8228, -3.4544, 4.2772]]
    gulp_eta = [0.0391, 0.0834, 0.0685, 0.0969, 0.1155, 0.0425, 0.1267, 0.0107, 0.6153]
    json.dump({"twocharges": _traj(f"{REF}/ex4/TRAJFF.twocharges"),
               "twocharges_expected": {"V1xx": 2.0003749, "V1yy": -1.0001874, "alpha": 0.8,
                                       "kmax": 11, "cutlongrange": 1000.0},
               "cluster": _traj(f"{REF}/ex4/TRAJFF.cluster"),
               "cluster_charges": [0.1, -0.8],
               "cluster_ewald": {"cutlongrange": 5.0, "epsw": 1e-8},
               "gulp_principal": gulp, "gulp_eta": gulp_eta}, open(f"{OUT}/ex4_efg.json", "w"), indent=1)


def _pot(path):
    vals = {}
    for l in open(path):
        l = l.split("!")[0].strip()
        if "=" not in l:
            continue
        k, v = l.split("=", 1)
        vals[k.strip()] = v.strip()
    return vals


def ex9():
    d = f"{REF}/ex9/config"
    t = open(f"{d}/POSFF").read().split("\n")
    n = int(t[0])
    A = np.array([[float(x) for x in t[k].split()[:3]] for k in (2, 3, 4)]).T
    names = t[6].split()
    rows = [l.split() for l in t[9:9 + n]]
    pos = np.array([[float(x) for x in r[1:4]] for r in rows])
    typ = np.array([names.index(r[0]) + 1 for r in rows], np.int32)
    pots = {}
    for f in ("IONS.POT", "PIMD.POT", "BMHFTD.POT"):
        pots.update(_pot(f"{d}/pot_mdff/{f}"))
    np.savez(f"{OUT}/ex9_pim.npz", pos=pos, itype=typ, A=A)
    json.dump({"names": names, "potentials": pots, "rcut": 7.2, "epsw": 1e-12,
               "conv_scf": 1e-6, "min_scf": 6, "max_scf": 100,
               "reference_utot": -8239.44074927, "reference_nscf": 14,
               "reference_alpha": 0.67659, "reference_kmax": 15},
              open(f"{OUT}/ex9_pim.json", "w"), indent=1)


if __name__ == "__main__":
    ex0()
    ex2()
    ex4()
    ex9()
