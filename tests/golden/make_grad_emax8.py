"""Write grad_emax8.json: dE0/dT of the full Magnus flow at eMax 8 from the
oracle's reverse sweep (CPU reference gradient), sampled at seeded elements.

Workload: 16O in the eMax 8 HO basis (hk8 orbits), synthetic vacuum
Hamiltonian seed 0, hw 20 MeV, normal ordered on the way in, white
generator, |mbpt2| < 1e-6 (the reference's single-precision literal), Omega
checkpointed every 4 steps.  The gradient is folded onto the independent
hermitian elements (adjoint.symmetric_gradient); kept are the 200 elements
of largest |g| and 2000 seeded elements with |g| > 1e-4.  Takes about half
an hour on 8 cores.

    python tests/golden/make_grad_emax8.py [threads]
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
from eom_imsrg.adjoint import symmetric_gradient  # noqa: E402
from oracle.oracle import OracleBasis, lib  # noqa: E402


def main():
    nthr = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    lib().orc_set_threads(nthr)
    b = Basis.oxygen16(8)
    o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
    T = synth.hamiltonian_vacuum(b, seed=0, hw=20.0)
    tol = float(np.float32(1e-6))
    t0 = time.time()
    e0, Tb, st = o.flow_grad(T, 0.5, tol, tol, tol, ckpt=4)
    dt = time.time() - t0
    g = symmetric_gradient(b, Tb, 1)
    order = np.argsort(-np.abs(g))
    top = order[:200]
    rng = np.random.default_rng(2024)
    cand = np.setdiff1d(np.nonzero(np.abs(g) > 1e-4)[0], top)
    pick = np.sort(rng.choice(cand, size=min(2000, len(cand)), replace=False))
    idx = np.concatenate([np.sort(top), pick])
    out = {"workload": "magnus_decouple_emax8 dE0/dT, seed 0, hw 20, white, |mbpt2| < 1e-6, ckpt 4",
           "E0": e0, "steps": int(st), "oplen": int(b.oplen), "threads": nthr, "seconds": round(dt, 1),
           "norm": float(np.linalg.norm(g)), "nonzero": int(np.count_nonzero(g)),
           "index": [int(x) for x in idx], "grad": [float(g[x]) for x in idx]}
    with open(os.path.join(HERE, "grad_emax8.json"), "w") as fh:
        json.dump(out, fh)
    print(f"E0 {e0} steps {st} in {dt:.0f} s; {len(idx)} elements written")


if __name__ == "__main__":
    main()
