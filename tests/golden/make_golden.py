"""Writes the committed fixtures of tests/golden/ (run in a checkout that has
the reference mounted at /root/reference; nothing at test time reads it).

hk_orbits.json    orbit tables (n, l, 2j, 2tz) of sp_inputs/hk{4,6,8,10}.sps
                  filtered to 2n+l <= eMax, and the hole flags of the "8 8"
                  row of sp_inputs/hole_scheme_hk (data files of the reference)
oracle_emax4.json the CPU oracle's flow of the eMax-4 configuration (16O holes,
                  hw 20 MeV, seed 0, white generator, reference tolerances):
                  E0 per Magnus step, nested commutators per step, final E0,
                  plus one [Omega, H] commutator's checksums
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"


def read_sps(path, emax):
    rows = []
    for line in open(path):
        f = line.split()
        if len(f) < 6:
            continue
        n, l, jj, tz = int(f[1]), int(f[2]), int(f[3]), int(f[4])
        if 2 * n + l <= emax:
            rows.append([n, l, jj, tz])
    return rows


def hole_row(path, Z, N):
    for line in open(path):
        f = line.split()
        if len(f) >= 3 and int(f[0]) == Z and int(f[1]) == N:
            return [int(x) for x in f[2:]]
    raise KeyError((Z, N))


def main():
    out = {}
    for e in (4, 6, 8, 10):
        out[f"hk{e}"] = read_sps(os.path.join(REF, "sp_inputs", f"hk{e}.sps"), e)
    out["holes_8_8"] = hole_row(os.path.join(REF, "sp_inputs", "hole_scheme_hk"), 8, 8)
    json.dump(out, open(os.path.join(HERE, "hk_orbits.json"), "w"))

    sys.path.insert(0, os.path.join(ROOT, "eom-imsrg_amd"))
    sys.path.insert(0, ROOT)
    from eom_imsrg import Basis, synth
    from oracle.oracle import OracleBasis
    ref_tol = float(np.float32(1e-6))
    b = Basis.oxygen16(4)
    o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    G = synth.antihermitian(b, seed=1, sigma=0.05)
    C = o.commutator(G, -1, H, 1, 1)
    HS, Om, log, st = o.magnus_decouple(H, 0.5, ref_tol, ref_tol, 0, ref_tol)
    gold = {
        "config": "eMax 4, 16O holes, hw 20 MeV, seed 0, white generator, ds 0.5, tolerances 1e-6 (single)",
        "E0_bare_normal_ordered": H[0],
        "mbpt2_bare": o.mbpt2(H),
        "commutator_Omega_H": {"E0": C[0], "sum": float(C.sum()), "sumsq": float((C * C).sum()),
                               "norm": o.frob_norm(C)},
        "flow": {"steps": int(st), "E0": log[:, 1].tolist(), "E0_plus_mbpt2": log[:, 2].tolist(),
                 "commutators": [int(x) for x in log[:, 3]], "final_E0": HS[0]},
    }
    json.dump(gold, open(os.path.join(HERE, "oracle_emax4.json"), "w"), indent=1)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
