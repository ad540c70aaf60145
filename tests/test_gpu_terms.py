"""HIP commutator terms vs the CPU oracle on the same seeded operators.

Every reference routine of the hot path is called through the C-ABI
(libimsrg_hip.so) on the GPU and through the oracle on the CPU, starting
from identical inputs (including a random prior RES, so the assign /
accumulate / symmetric-fill semantics are exercised).  Tolerance: FP64,
1e-10 relative per element (see conftest.assert_rel for the floor applied
to elements cancelled far below their block's scale).
"""
import numpy as np
import pytest
import torch

from conftest import assert_rel, basis, oracle_basis

pytestmark = pytest.mark.gpu

HERMS = [(-1, 1), (-1, -1), (1, 1), (1, -1)]


def _ops(b, hL, hR, seed=11):
    from eom_imsrg import synth
    mk = lambda h, s: synth.hermitian_random(b, s, 1.0) if h == 1 else synth.antihermitian(b, s, 1.0)
    L = mk(hL, seed)
    R = mk(hR, seed + 1)
    prior = synth.hermitian_random(b, seed + 2, 0.3)
    prior += np.random.default_rng(seed).normal(0, 0.1, prior.shape)   # no symmetry at all
    return L, R, prior


def _dev(b, arr, herm):
    from eom_imsrg import SqOp
    return SqOp.from_numpy(b, arr, herm)


def _sync():
    torch.cuda.synchronize()


@pytest.fixture(params=[0, 1], ids=["grouped", "plain"])
def path(request):
    return request.param


@pytest.mark.parametrize("emax", [2, 4])
@pytest.mark.parametrize("hL,hR", HERMS)
def test_terms_vs_oracle(gpu, emax, hL, hR, path):
    from eom_imsrg import commutators as cm, SqOp, CCMat, CCWork
    b, o = basis(emax), oracle_basis(emax)
    b.set_path(path)
    print(f"[terms] emax={emax} herm=({hL},{hR}) path={path}", flush=True)
    L, R, prior = _ops(b, hL, hR)
    hres = -hL * hR
    Ld, Rd = _dev(b, L, hL), _dev(b, R, hR)

    # 110 / 220
    assert abs(cm.commutator_110(Ld, Rd) - o.comm110(L, hL, R, hR)) <= 1e-10 * (1 + abs(o.comm110(L, hL, R, hR)))
    assert abs(cm.commutator_220(Ld, Rd) - o.comm220(L, hL, R, hR)) <= 1e-10 * (1 + abs(o.comm220(L, hL, R, hR)))

    for name in ("111", "121", "122"):
        ref = prior.copy()
        getattr(o, "comm" + name)(L, hL, R, hR, ref, hres)
        RES = _dev(b, prior, hres)
        getattr(cm, "commutator_" + name)(Ld, Rd, RES)
        _sync()
        assert_rel(RES.numpy(), ref, what=f"commutator_{name}")

    # 222_pp_hh with its w1 / w2 outputs
    ref = prior.copy()
    w1r, w2r = np.zeros(b.oplen), np.zeros(b.oplen)
    o.comm222pphh(L, hL, R, hR, ref, hres, w1r, w2r)
    RES = _dev(b, prior, hres)
    w1, w2 = SqOp.zeros(b), SqOp.zeros(b)
    cm.commutator_222_pp_hh(Ld, Rd, RES, w1, w2)
    _sync()
    assert_rel(RES.numpy(), ref, what="commutator_222_pp_hh RES")
    two = slice(b.off_fpp + b.npart * b.npart, b.oplen)
    assert_rel(w1.numpy()[two], w1r[two], what="222_pp_hh w1")
    assert_rel(w2.numpy()[two], w2r[two], what="222_pp_hh w2")

    # 221 on the oracle's w1 / w2
    ref = prior.copy()
    o.comm221(L, hL, R, hR, ref, hres, w1r, w2r)
    RES = _dev(b, prior, hres)
    cm.commutator_221(Ld, Rd, RES, _dev(b, w1r, 1), _dev(b, w2r, 1))
    _sync()
    assert_rel(RES.numpy(), ref, what="commutator_221")

    # Pandya transform of both operands
    LCCr, RCCr = o.cross_coupled(L, hL), o.cross_coupled(R, hR)
    LCC, RCC = CCMat(b, hL), CCMat(b, hR)
    cm.calculate_cross_coupled(Ld, LCC)
    cm.calculate_cross_coupled(Rd, RCC)
    _sync()
    assert_rel(LCC.data.cpu().numpy(), LCCr, what="calculate_cross_coupled L")
    assert_rel(RCC.data.cpu().numpy(), RCCr, what="calculate_cross_coupled R")

    # 222_ph on the oracle's cross-coupled matrices
    ref = prior.copy()
    Wr = o.comm222ph(LCCr, hL, RCCr, hR, ref, hres)
    RES = _dev(b, prior, hres)
    W = CCWork(b)
    LCCd, RCCd = CCMat(b, hL), CCMat(b, hR)
    LCCd.data.copy_(torch.from_numpy(LCCr))
    RCCd.data.copy_(torch.from_numpy(RCCr))
    cm.commutator_222_ph(LCCd, RCCd, RES, W)
    _sync()
    # WCC holds W' = Lh W - Rh W^T (upper triangle of every block)
    Wg = W.data.cpu().numpy()
    wscale = 2 * float(np.max(np.abs(Wr)))
    for q in range(b.ncc):
        r, off = int(b.rlen[q]), int(b.wcoff[q])
        if r == 0:
            continue
        Wq = Wr[off: off + r * r].reshape(r, r, order="F")
        Wp = hL * Wq - hR * Wq.T
        iu = np.triu_indices(r)
        assert_rel(Wg[off: off + r * r].reshape(r, r, order="F")[iu], Wp[iu], what=f"222_ph WCC block {q}",
                   scale=wscale)
    assert_rel(RES.numpy(), ref, what="commutator_222_ph RES")
    b.set_path(0)


@pytest.mark.parametrize("emax", [2, 4, 6])
@pytest.mark.parametrize("hL,hR", [(-1, 1), (-1, -1)])
def test_full_commutator(gpu, emax, hL, hR, path):
    """RES = [L, R] in the BCH term order (IMSRG_MAGNUS.f90:207-219)"""
    from eom_imsrg import commutators as cm, SqOp
    b, o = basis(emax), oracle_basis(emax)
    b.set_path(path)
    print(f"[full] emax={emax} herm=({hL},{hR}) path={path}", flush=True)
    L, R, _ = _ops(b, hL, hR, seed=21)
    hres = -hL * hR
    ref = o.commutator(L, hL, R, hR, hres)
    RES = SqOp.zeros(b, hres)
    cm.commutator(_dev(b, L, hL), _dev(b, R, hR), RES)
    _sync()
    assert_rel(RES.numpy(), ref, what=f"[L,R] emax {emax}")
    b.set_path(0)


@pytest.mark.parametrize("emax", [2, 4])
def test_algebra_generator_mbpt2(gpu, emax):
    from eom_imsrg import commutators as cm, SqOp, synth
    b, o = basis(emax), oracle_basis(emax)
    print(f"[algebra] emax={emax}", flush=True)
    Hv = synth.hamiltonian_vacuum(b, seed=0, hw=20.0)
    Href = o.normal_order(Hv)
    H = _dev(b, Hv, 1)
    cm.normal_order(H)
    _sync()
    assert_rel(H.numpy(), Href, what="normal_order")
    eta = SqOp.zeros(b, -1)
    cm.build_gs_white(_dev(b, Href, 1), eta)
    _sync()
    assert_rel(eta.numpy()[1:], o.white(Href)[1:], what="build_gs_white")
    assert abs(cm.mbpt2(_dev(b, Href, 1)) - o.mbpt2(Href)) <= 1e-10 * abs(o.mbpt2(Href))
    assert abs(cm.mat_frob_norm(_dev(b, Href, 1)) - o.frob_norm(Href)) <= 1e-12 * o.frob_norm(Href)
    A = _dev(b, Href, 1)
    Bm = _dev(b, o.white(Href), -1)
    C = SqOp.zeros(b)
    cm.add_sq_op(A, 1.0, Bm, 0.37, C)
    _sync()
    assert_rel(C.numpy(), o.add(Href, 1.0, o.white(Href), 0.37), rtol=1e-15, what="add_sq_op")
