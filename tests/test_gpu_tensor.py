"""GPU parity of the tensor path (TS_commutator_* / BCH_TENSOR) against the
CPU oracle, term by term, at eMax 2-4 for ranks 0-6 with and without parity
change; at eMax 6 and 8 through properties the oracle does not need: the
rank-0 tensor commutator equals the (separately verified) scalar GPU
commutator times sqrt(2J+1), and the commutator is linear in the tensor.
Tolerance: 1e-10 of the largest element of each compared array (the 222_ph
cut |W| > 1e-6 is applied identically; the scalar 222_ph's .70710678118
literal differs from the tensor's sqrt(2) by 1.9e-11 relative)."""
import math
import os
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "eom-imsrg_amd"))
sys.path.insert(0, ROOT)

from eom_imsrg import Basis, SqOp, CCMat, synth, tensor, commutators as cm  # noqa: E402
from oracle.oracle import OracleBasis, OracleTensor, use_host_blas  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _close(got, ref, what, tol=TOL):
    got, ref = np.asarray(got), np.asarray(ref)
    sc = max(float(np.max(np.abs(ref))), 1e-300)
    err = float(np.max(np.abs(got - ref))) if ref.size else 0.0
    assert err <= tol * sc, f"{what}: max err {err:.3e} (scale {sc:.3e})"


def _case(emax, rank, dpar, hL, hR, seed=0):
    use_host_blas()
    b = Basis.oxygen16(emax)
    o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
    ot = OracleTensor(o, rank, dpar)
    tl = tensor.TensorLayout(b, rank, dpar)
    L = synth.antihermitian(b, seed=11 + seed, sigma=0.3) if hL < 0 else synth.hermitian_random(b, seed=11 + seed,
                                                                                                   sigma=0.5)
    R = synth.tensor_random(tl, seed=12 + seed, herm=hR)
    P = synth.tensor_random(tl, seed=13 + seed, herm=-hL * hR)   # prior RES: checks assign / accumulate
    return b, o, ot, tl, L, R, P


CASES = [(2, 4, 0, -1, 1), (2, 2, 2, -1, -1), (2, 0, 0, -1, 1), (3, 6, 2, -1, 1), (4, 4, 0, -1, 1),
         (4, 2, 2, 1, -1)]


@pytest.mark.parametrize("emax,rank,dpar,hL,hR", CASES)
def test_terms_vs_oracle(emax, rank, dpar, hL, hR):
    b, o, ot, tl, L, R, P = _case(emax, rank, dpar, hL, hR)
    hC = -hL * hR
    print(f"tensor terms eMax {emax} rank {rank} dpar {dpar} herm ({hL},{hR})", flush=True)
    Ld = SqOp.from_numpy(b, L, hL)
    Rd = tensor.TensorOp.from_numpy(tl, R, hR)
    n1 = b.off_fpp + b.npart * b.npart      # one-body part [0, n1)

    def run(gpu_fn, orc_fn, what, part=slice(None)):
        RES = tensor.TensorOp.from_numpy(tl, P, hC)
        gpu_fn(RES)
        torch.cuda.synchronize()
        ref = orc_fn(P.copy())
        _close(RES.numpy()[part], ref[part], what)
        return RES

    run(lambda X: tensor.TS_commutator_111(Ld, Rd, X), lambda X: ot.ts111(L, hL, R, hR, X, hC), "111",
        slice(0, n1))
    run(lambda X: tensor.TS_commutator_121(Ld, Rd, X), lambda X: ot.ts121(L, hL, R, hR, X, hC), "121")
    LCC = CCMat(b, hL)
    cm.calculate_cross_coupled(Ld, LCC)
    run(lambda X: tensor.TS_commutator_211(LCC, Rd, X), lambda X: ot.ts211(L, hL, R, hR, X, hC), "211")
    run(lambda X: tensor.TS_commutator_122(Ld, Rd, X), lambda X: ot.ts122(L, hL, R, hR, X, hC), "122")
    run(lambda X: tensor.TS_commutator_212(Ld, Rd, X), lambda X: ot.ts212(L, hL, R, hR, X, hC), "212")
    w1, w2 = tensor.TensorOp.zeros(tl, hR), tensor.TensorOp.zeros(tl, hR)
    ow = {}

    def orc_pphh(X):
        Y, a1, a2 = ot.ts222pphh(L, hL, R, hR, X, hC)
        ow["w1"], ow["w2"] = a1, a2
        return Y

    run(lambda X: tensor.TS_commutator_222_pp_hh(Ld, Rd, X, w1, w2), orc_pphh, "222_pp_hh")
    two = slice(n1, tl.oplen)
    _close(w1.numpy()[two], ow["w1"][two], "222_pp_hh w1")
    _close(w2.numpy()[two], ow["w2"][two], "222_pp_hh w2")
    W1 = tensor.TensorOp.from_numpy(tl, ow["w1"], 1)
    W2 = tensor.TensorOp.from_numpy(tl, ow["w2"], 1)
    run(lambda X: tensor.TS_commutator_221(W1, W2, hL * hR, X), lambda X: ot.ts221(ow["w1"], ow["w2"], hL * hR, X, hC),
        "221")
    RCC = tensor.TensorCC(tl)
    tensor.calculate_single_pandya(Rd, RCC)
    torch.cuda.synchronize()
    _close(RCC.data.cpu().numpy()[: tl.cclen], ot.pandya(R, hR), "single_pandya")
    run(lambda X: tensor.TS_commutator_222_ph(LCC, RCC, Rd, X), lambda X: ot.ts222ph(L, hL, R, hR, X, hC), "222_ph")


@pytest.mark.parametrize("emax,rank,dpar,hL,hR", CASES)
def test_full_commutator_and_norm_vs_oracle(emax, rank, dpar, hL, hR):
    b, o, ot, tl, L, R, P = _case(emax, rank, dpar, hL, hR, seed=3)
    hC = -hL * hR
    print(f"tensor commutator eMax {emax} rank {rank} dpar {dpar}", flush=True)
    RES = tensor.TensorOp.zeros(tl, hC)
    tensor.TS_commutator(SqOp.from_numpy(b, L, hL), tensor.TensorOp.from_numpy(tl, R, hR), RES)
    torch.cuda.synchronize()
    ref = ot.commutator(L, hL, R, hR, hC)
    _close(RES.numpy(), ref, "[L, R]")
    n = tensor.mat_frob_norm(RES)
    assert abs(n - ot.frob_norm(ref)) <= 1e-12 * ot.frob_norm(ref)
    # bit-reproducible: a second call gives the same bytes
    RES2 = tensor.TensorOp.zeros(tl, hC)
    tensor.TS_commutator(SqOp.from_numpy(b, L, hL), tensor.TensorOp.from_numpy(tl, R, hR), RES2)
    assert torch.equal(RES.data, RES2.data)


@pytest.mark.parametrize("emax,rank,dpar", [(2, 4, 0), (3, 2, 2)])
def test_bch_tensor_vs_oracle(emax, rank, dpar):
    b, o, ot, tl, L, R, P = _case(emax, rank, dpar, -1, 1)
    print(f"BCH_TENSOR eMax {emax} rank {rank}", flush=True)
    G = synth.antihermitian(b, seed=21, sigma=0.1)
    HS = tensor.TensorOp.from_numpy(tl, R, 1)
    it, adv = tensor.BCH_TENSOR(SqOp.from_numpy(b, G, -1), HS)
    ref, itr, advr = ot.bch(G, -1, R, 1)
    assert it == itr
    _close(HS.numpy(), ref, "BCH_TENSOR")
    _close(adv[:it], advr[:itr], "advals", 1e-9)


def test_bch_tensor_zero_generator():
    b, o, ot, tl, L, R, P = _case(2, 4, 0, -1, 1)
    HS = tensor.TensorOp.from_numpy(tl, R, 1)
    it, adv = tensor.BCH_TENSOR(SqOp.zeros(b, -1), HS)
    assert it == 2 and adv[1] == 0.0
    assert np.array_equal(HS.numpy(), R)


def _rank0(b, tl, S, herm):
    """copy_rank0_to_tensor_format (basic_IMSRG.f90:3037-3074) in numpy"""
    T = np.zeros(tl.oplen)
    A, Pn = b.nhole, b.npart
    jh = np.sqrt(np.asarray(b.jj)[np.asarray(b.holes)] + 1.0)
    jp = np.sqrt(np.asarray(b.jj)[np.asarray(b.parts)] + 1.0)
    fhh = S[b.off_fhh: b.off_fhh + A * A].reshape(A, A, order="F") * jh[:, None]
    fph = S[b.off_fph: b.off_fph + Pn * A].reshape(Pn, A, order="F") * jp[:, None]
    fpp = S[b.off_fpp: b.off_fpp + Pn * Pn].reshape(Pn, Pn, order="F") * jp[:, None]
    T[b.off_fhh: b.off_fhh + A * A] = fhh.reshape(-1, order="F")
    T[b.off_fph: b.off_fph + Pn * A] = fph.reshape(-1, order="F")
    T[b.off_fpp: b.off_fpp + Pn * Pn] = fpp.reshape(-1, order="F")
    look = {tuple(int(x) for x in tl.info[q, :4]): q for q in range(tl.nblocks)}
    src = (0, 1, 2, 3, 4, 5, 2, 1, 5)
    for q in range(b.nblocks):
        J, Tz, Pr = (int(x) for x in b.jtp[q])
        qT = look[(J, J, Tz, Pr)]
        s = math.sqrt(J + 1.0)
        for g in range(9):
            n1, n2 = tl.shape(qT, g)
            if n1 * n2 == 0:
                continue
            r, c = b.block_shape(q, src[g])
            X = S[int(b.goff[q, src[g]]): int(b.goff[q, src[g]]) + r * c].reshape(r, c, order="F")
            Y = X * s if g < 6 else herm * X.T * s
            o = int(tl.goff[qT, g])
            T[o: o + n1 * n2] = Y.reshape(-1, order="F")
    return T


@pytest.mark.parametrize("emax", [6, 8])
def test_rank0_tensor_equals_scalar_path(emax):
    """at the configurations' sizes: [Omega, S] through the tensor kernels
    (rank 0) equals the scalar grouped commutator, element by element.  S is
    scaled by 1e4 so that no cross-coupled product of the tensor 222_ph
    falls under its |W| > 1e-6 cut (the scalar path has none): at unit
    scale the dropped products move elements by ~1e-6."""
    b = Basis.oxygen16(emax)
    tl = tensor.TensorLayout(b, 0, 0)
    print(f"rank-0 tensor vs scalar eMax {emax}", flush=True)
    G = synth.antihermitian(b, seed=31, sigma=0.05)
    S = 1e4 * synth.hermitian_random(b, seed=32)
    C = SqOp.zeros(b, 1)
    cm.commutator(SqOp.from_numpy(b, G, -1), SqOp.from_numpy(b, S, 1), C)
    RES = tensor.TensorOp.zeros(tl, 1)
    tensor.TS_commutator(SqOp.from_numpy(b, G, -1), tensor.TensorOp.from_numpy(tl, _rank0(b, tl, S, 1), 1), RES)
    torch.cuda.synchronize()
    ref = _rank0(b, tl, C.numpy(), 1)
    got = RES.numpy()
    _close(got[1:], ref[1:], f"rank-0 tensor vs scalar eMax {emax}")


def test_linearity_emax8():
    b = Basis.oxygen16(8)
    tl = tensor.TensorLayout(b, 4, 0)
    print("tensor commutator linearity eMax 8 rank 4", flush=True)
    G = SqOp.from_numpy(b, synth.antihermitian(b, seed=41, sigma=0.05), -1)
    # scaled as above, so the 222_ph cut |W| > 1e-6 drops the same (no) products in all three calls
    R1, R2 = 1e4 * synth.tensor_random(tl, seed=42, herm=1), 1e4 * synth.tensor_random(tl, seed=43, herm=1)
    out = []
    for X in (R1, R2, R1 + 2.0 * R2):
        RES = tensor.TensorOp.zeros(tl, 1)
        tensor.TS_commutator(G, tensor.TensorOp.from_numpy(tl, X, 1), RES)
        out.append(RES.numpy())
    _close(out[2], out[0] + 2.0 * out[1], "linearity", 1e-12)


def _fixture(name):
    import json
    return json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)))


def _check_sample(tl, b, X, fx, what):
    n1 = b.off_fpp + b.npart * b.npart
    _close(X[:n1], np.asarray(fx["one_body"]), what + " one-body")
    _close(X[np.asarray(fx["idx"])], np.asarray(fx["val"]), what + " two-body sample")
    n = tensor.mat_frob_norm(tensor.TensorOp.from_numpy(tl, X, 1))
    assert abs(n - fx["norm"]) <= 1e-12 * fx["norm"], (what, n, fx["norm"])


@pytest.mark.parametrize("rank,dpar", [(4, 0), (2, 2)])
def test_commutator_emax6_fixture(rank, dpar):
    """[L, R] at eMax 6 against the oracle's (tests/golden/make_tensor_fixtures.py):
    one-body part, 4,096 seeded two-body elements, and the norm of the whole result"""
    fx = _fixture("tensor_emax6.json")[f"rank{rank}_dpar{dpar}"]
    b = Basis.oxygen16(6)
    tl = tensor.TensorLayout(b, rank, dpar)
    print(f"tensor commutator eMax 6 rank {rank} dpar {dpar} vs fixture", flush=True)
    L = SqOp.from_numpy(b, synth.antihermitian(b, seed=11, sigma=0.3), -1)
    R = tensor.TensorOp.from_numpy(tl, synth.tensor_random(tl, seed=12, herm=1), 1)
    RES = tensor.TensorOp.zeros(tl, 1)
    tensor.TS_commutator(L, R, RES)
    torch.cuda.synchronize()
    _check_sample(tl, b, RES.numpy(), fx, f"eMax 6 rank {rank}")


def test_bch_tensor_emax4_fixture():
    fx = _fixture("tensor_bch_emax4.json")
    b = Basis.oxygen16(4)
    tl = tensor.TensorLayout(b, 4, 0)
    print("BCH_TENSOR eMax 4 vs fixture", flush=True)
    G = SqOp.from_numpy(b, synth.antihermitian(b, seed=21, sigma=0.1), -1)
    HS = tensor.TensorOp.from_numpy(tl, synth.tensor_random(tl, seed=22, herm=1), 1)
    it, adv = tensor.BCH_TENSOR(G, HS)
    assert it == fx["iterations"]
    _close(adv[:it], np.asarray(fx["advals"]), "advals", 1e-9)
    _check_sample(tl, b, HS.numpy(), fx, "BCH_TENSOR eMax 4")
