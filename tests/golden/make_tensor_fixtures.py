"""Write the oracle fixtures of the larger tensor GPU tests (CPU oracle, run
once in the development container; the GPU tests read the JSON only).

tensor_emax6.json   eMax 6 (16O holes): RES = [L, R] (TS_commutator
                    sequence) for L = synth.antihermitian(seed 11, sigma
                    0.3) and R = synth.tensor_random(seed 12, herm +1) of
                    rank 4 (even parity) and of rank 2 with parity change:
                    the whole one-body part, 4096 seeded two-body elements
                    and the Frobenius norm of RES (mat_frob_norm's weights).
tensor_bch_emax4.json  eMax 4: BCH_TENSOR of R (rank 4, seed 22) with
                    G = synth.antihermitian(seed 21, sigma 0.1): last term,
                    ||term||*coef per term, the one-body part, 4096 seeded
                    elements and the norm of the result.

    python tests/golden/make_tensor_fixtures.py [emax6|bch ...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "eom-imsrg_amd"), ROOT]

from eom_imsrg import Basis, synth, tensor  # noqa: E402
from oracle.oracle import OracleBasis, OracleTensor, use_host_blas  # noqa: E402


def _sample(tl, b, X, n=4096, seed=7):
    n1 = b.off_fpp + b.npart * b.npart
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(np.arange(n1, tl.oplen), size=min(n, tl.oplen - n1), replace=False))
    return {"one_body": X[:n1].tolist(), "idx": idx.tolist(), "val": X[idx].tolist()}


def emax6():
    b = Basis.oxygen16(6)
    o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
    out = {}
    for rank, dpar in ((4, 0), (2, 2)):
        t0 = time.time()
        ot = OracleTensor(o, rank, dpar)
        tl = tensor.TensorLayout(b, rank, dpar)
        L = synth.antihermitian(b, seed=11, sigma=0.3)
        R = synth.tensor_random(tl, seed=12, herm=1)
        C = ot.commutator(L, -1, R, 1, 1)
        out[f"rank{rank}_dpar{dpar}"] = {"norm": ot.frob_norm(C), **_sample(tl, b, C)}
        print(f"eMax 6 rank {rank} dpar {dpar}: {time.time() - t0:.0f} s", flush=True)
    json.dump(out, open(os.path.join(HERE, "tensor_emax6.json"), "w"))


def bch():
    b = Basis.oxygen16(4)
    o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
    ot = OracleTensor(o, 4, 0)
    tl = tensor.TensorLayout(b, 4, 0)
    G = synth.antihermitian(b, seed=21, sigma=0.1)
    R = synth.tensor_random(tl, seed=22, herm=1)
    t0 = time.time()
    out, it, adv = ot.bch(G, -1, R, 1)
    print(f"eMax 4 BCH_TENSOR: {it} terms, {time.time() - t0:.0f} s", flush=True)
    json.dump({"iterations": it, "advals": adv[:it].tolist(), "norm": ot.frob_norm(out), **_sample(tl, b, out)},
              open(os.path.join(HERE, "tensor_bch_emax4.json"), "w"))


if __name__ == "__main__":
    use_host_blas()
    which = sys.argv[1:] or ["emax6", "bch"]
    for w in which:
        {"emax6": emax6, "bch": bch}[w]()
