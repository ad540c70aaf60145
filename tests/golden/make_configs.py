"""Write the oracle fixtures of the configuration-size GPU tests (CPU oracle,
run once in the development container; the GPU tests read the JSON only).

configs_emax10.json   eMax 10 (hk10, 132 orbits, 16O holes), H = normal
                      ordered vacuum Hamiltonian seed 0, hw 20: the first
                      Magnus step (Omega_1 = 0.5 * white(H), BCH terms to
                      1e-10) with its BCH exit index, ||term||*coef per term,
                      E0, mbpt2 and 4096 seeded elements of H(s = 0.5); then
                      MAGNUS_EXPAND of the second step (its nested-commutator
                      count and the norm of dOmega).
configs_emax8_flow.json  eMax 8, the same Hamiltonian recipe, white
                      generator, flowed until |E0(s) - E0(s - ds)| < 1e-8 MeV:
                      E0 and commutators per step.
configs_batch_emax6.json  eMax 6, the batch instances 0..5 (seed k, two-body
                      scale U(0.8, 1.2), hw in {16, 20, 24}), each flowed to
                      |mbpt2| < 1e-6: E0 and steps per instance.

    python tests/golden/make_configs.py [threads] [emax10|emax8|batch ...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "eom-imsrg_amd"), ROOT]

from eom_imsrg import Basis, synth  # noqa: E402
from oracle.oracle import OracleBasis, lib  # noqa: E402

REF = float(np.float32(1e-6))


def _oracle(emax):
    b = Basis.oxygen16(emax)
    return b, OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)


def emax10():
    b, o = _oracle(10)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    t0 = time.time()
    eta = o.white(H)
    G1 = 0.5 * eta
    HS, nb, adv = o.bch(G1, H, 1e-10)
    m = o.mbpt2(HS)
    DG, nm = o.magnus(o.white(HS), G1, 1e-10)
    rng = np.random.default_rng(10)
    idx = np.sort(rng.choice(b.oplen, 4096, replace=False))
    idx[0] = 0
    out = {"workload": "eMax 10 first Magnus step (Omega_1 = 0.5 white(H)), BCH to 1e-10, seed 0, hw 20",
           "E0_bare": float(H[0]), "bch_index": int(nb), "advals": [float(x) for x in adv[: max(nb - 1, 0)]],
           "E0": float(HS[0]), "mbpt2": float(m), "norm": float(o.frob_norm(HS)),
           "index": [int(x) for x in idx], "HS": [float(HS[x]) for x in idx],
           "magnus2_commutators": int(nm), "magnus2_norm": float(o.frob_norm(DG)),
           "seconds": round(time.time() - t0, 1)}
    json.dump(out, open(os.path.join(HERE, "configs_emax10.json"), "w"))
    print("emax10", out["bch_index"], out["E0"], out["magnus2_commutators"], out["seconds"], flush=True)


def emax8():
    b, o = _oracle(8)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    t0 = time.time()
    HS, _, log, st = o.magnus_decouple(H, 0.5, REF, REF, 1, 1e-8, maxsteps=100)
    out = {"workload": "eMax 8 magnus_decouple, seed 0, hw 20, white, until |dE0| < 1e-8 MeV",
           "steps": int(st), "E0": [float(x) for x in log[:, 1]], "commutators": [int(x) for x in log[:, 3]],
           "mbpt2": [float(x - y) for x, y in zip(log[:, 2], log[:, 1])], "seconds": round(time.time() - t0, 1)}
    json.dump(out, open(os.path.join(HERE, "configs_emax8_flow.json"), "w"))
    print("emax8", st, out["E0"][-1], out["seconds"], flush=True)


def batch():
    b, o = _oracle(6)
    rows = []
    t0 = time.time()
    for k in range(6):
        seed, scale, hw = synth.instance_parameters(k)
        H = o.normal_order(synth.hamiltonian_vacuum(b, seed=seed, hw=hw, v_scale=scale))
        HS, _, log, st = o.magnus_decouple(H, 0.5, REF, REF, 0, REF)
        rows.append({"instance": k, "hw": hw, "scale": scale, "steps": int(st), "E0": float(HS[0]),
                     "E0_bare": float(H[0])})
        print("batch", rows[-1], flush=True)
    out = {"workload": "eMax 6 batch instances 0..5, white, |mbpt2| < 1e-6", "instances": rows,
           "seconds": round(time.time() - t0, 1)}
    json.dump(out, open(os.path.join(HERE, "configs_batch_emax6.json"), "w"))


def main():
    nthr = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    lib().orc_set_threads(nthr)
    which = sys.argv[2:] or ["batch", "emax10", "emax8"]
    for w in which:
        globals()[w]()


if __name__ == "__main__":
    main()
