"""The generator family on the GPU vs the oracle: build_gs_atan,
build_gs_atanxx, build_gs_w2 (generators.f90:109-374) per element at
1e-10 relative, and magnus_decouple driven by each generator (single flow
and batch) against the oracle's flow with the same generator: equal step
and commutator counts, E0 within 1e-8 MeV."""
import numpy as np
import pytest
import torch

from conftest import assert_rel, basis, oracle_basis

pytestmark = pytest.mark.gpu

REF = float(np.float32(1e-6))


def _hamiltonian(emax, seed=0, hw=20.0, scale=1.0):
    from eom_imsrg import synth
    b, o = basis(emax), oracle_basis(emax)
    return o.normal_order(synth.hamiltonian_vacuum(b, seed=seed, hw=hw, v_scale=scale))


def test_atan_atanxx_w2_vs_oracle(gpu):
    from eom_imsrg import Flow, SqOp, commutators as cm, synth
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    print("[generators] emax=4 atan / atanxx / w2", flush=True)
    H = _hamiltonian(emax)
    # a partly flowed H as the numerator source of atanxx
    H2 = o.commutator(synth.antihermitian(b, 7, 0.05), -1, H, 1, 1)
    Hd, H2d = SqOp.from_numpy(b, H, 1), SqOp.from_numpy(b, H2, 1)
    eta = SqOp.zeros(b, -1)
    cm.build_gs_atan(Hd, eta)
    torch.cuda.synchronize()
    assert_rel(eta.numpy()[1:], o.gs_atan(H)[1:], what="build_gs_atan")
    eta2 = SqOp.zeros(b, -1)
    cm.build_gs_atanxx(Hd, H2d, eta2)
    torch.cuda.synchronize()
    assert_rel(eta2.numpy()[1:], o.gs_atanxx(H, H2)[1:], what="build_gs_atanxx")
    f = Flow(b)
    w = SqOp.zeros(b, -1)
    f.build_gs_w2(Hd, w)
    torch.cuda.synchronize()
    ref = o.gs_w2(H)
    assert w.E0 == 0.0
    assert_rel(w.numpy(), ref, what="build_gs_w2")
    # the input operator is left as it was
    assert np.array_equal(Hd.numpy(), H)


@pytest.mark.parametrize("gen", ["atan", "w2"])
def test_magnus_decouple_with_generator(gpu, gen):
    from eom_imsrg import Flow, SqOp
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    gid = {"white": 0, "atan": 1, "w2": 2}[gen]
    print(f"[flow generator] emax=4 generator={gen}", flush=True)
    H = _hamiltonian(emax)
    HSr, Gr, logr, str_ = o.magnus_decouple(H, 0.5, REF, REF, 0, REF, generator=gid)
    f = Flow(b, graphs=True, generator=gen)
    HS = SqOp.from_numpy(b, H, 1)
    G = SqOp.zeros(b, -1)
    log = f.magnus_decouple(HS, G, ds=0.5)
    torch.cuda.synchronize()
    assert len(log) == str_
    assert np.array_equal(log[:, 3], logr[:, 3])
    assert abs(HS.E0 - HSr[0]) < 1e-8, (HS.E0, HSr[0])
    assert np.max(np.abs(log[:, 1] - logr[:, 1])) < 1e-8
    assert_rel(HS.numpy(), HSr, rtol=1e-9, what="decoupled H")


def test_batch_decouple_w2_rows(gpu):
    """a w2 batch of 3 instances: each row equals the single w2 flow, and the
    first against the oracle's w2 flow"""
    from eom_imsrg import Batch, Flow, SqOp, synth
    emax, ni = 4, 3
    b, o = basis(emax), oracle_basis(emax)
    print(f"[batch generator] emax=4 ninst={ni} w2", flush=True)
    Hs = []
    for k in range(ni):
        seed, scale, hw = synth.instance_parameters(k)
        Hs.append(_hamiltonian(emax, seed, hw, scale))
    H = np.stack(Hs)
    bt = Batch(b, ni, graphs=True, generator="w2")
    HS = torch.from_numpy(H.copy()).cuda()
    G = bt.zeros()
    e0, steps, nmax = bt.magnus_decouple(HS, G)
    torch.cuda.synchronize()
    f = Flow(b, graphs=True, generator="w2")
    for k in range(ni):
        hs = SqOp.from_numpy(b, H[k], 1)
        log = f.magnus_decouple(hs, SqOp.zeros(b, -1))
        torch.cuda.synchronize()
        assert len(log) == steps[k]
        assert abs(log[-1, 1] - e0[k]) <= 1e-12 * abs(e0[k])
        assert_rel(HS[k].cpu().numpy(), hs.numpy(), rtol=1e-12, what=f"w2 row {k}")
    HSr, _, _, str_ = o.magnus_decouple(H[0], 0.5, REF, REF, 0, REF, generator=2)
    assert str_ == steps[0] and abs(HSr[0] - e0[0]) < 1e-8


def test_unknown_generator_is_refused(gpu):
    from eom_imsrg import Flow
    f = Flow(basis(2))
    with pytest.raises(KeyError):
        f.set_generator("imtime")


@pytest.mark.parametrize("gen", ["atan", "w2"])
def test_generator_vjp_vs_oracle(gpu, gen):
    """transpose of build_gs_atan / build_gs_w2 against the oracle's, eMax 4"""
    from eom_imsrg import Flow, SqOp, adjoint
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    print(f"[generator vjp] emax=4 {gen}", flush=True)
    H = _hamiltonian(emax)
    Eb = np.random.default_rng(6).normal(size=b.oplen)
    if gen == "atan":
        Hdb, Hnb = o.gen_vjp(H, H, Eb, 1)
        ref = Hdb + Hnb
    else:
        ref = o.gs_w2_vjp(H, Eb)
    f = Flow(b, generator=gen)
    Hb = torch.zeros(b.oplen, dtype=torch.float64, device="cuda")
    adjoint.generator_vjp(f, SqOp.from_numpy(b, H, 1), torch.from_numpy(Eb).cuda(), Hb)
    torch.cuda.synchronize()
    assert_rel(adjoint.symmetric_gradient(b, Hb.cpu().numpy(), 1), adjoint.symmetric_gradient(b, ref, 1),
               rtol=1e-9, what=f"{gen} transpose")


@pytest.mark.parametrize("gen", ["atan", "w2"])
def test_flow_gradient_with_generator_vs_oracle(gpu, gen):
    """dE0/dT of the whole flow with the atan / w2 generator (Omega and H(s)
    checkpointed every 4 steps) against the oracle's reverse sweep, eMax 4"""
    from eom_imsrg import Flow, adjoint, synth
    b, o = basis(4), oracle_basis(4)
    print(f"[flow_grad generator] emax=4 {gen}", flush=True)
    T = synth.hamiltonian_vacuum(b, 0)
    gid = {"atan": 1, "w2": 2}[gen]
    e0_o, Tb_o, st_o = o.flow_grad(T, 0.5, REF, REF, REF, ckpt=4, generator=gid)
    e0, Tb, st = adjoint.flow_grad(Flow(b, graphs=True, generator=gen), torch.from_numpy(T).cuda(), ckpt=4)
    assert st == st_o and abs(e0 - e0_o) < 1e-8, (st, st_o, e0, e0_o)
    assert_rel(adjoint.symmetric_gradient(b, Tb.cpu().numpy(), 1), adjoint.symmetric_gradient(b, Tb_o, 1), rtol=1e-8,
               what=f"dE0/dT {gen}")


def test_transform_sq_op_bch(gpu):
    """transform_sq_op_BCH: the bare H transformed by a flow's final Omega is
    that flow's decoupled H (the flow's last BCH), and an observable against
    the oracle's BCH with equal term counts"""
    from eom_imsrg import Flow, SqOp, synth
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    print("[transform_sq_op_BCH] emax=4", flush=True)
    H = _hamiltonian(emax)
    f = Flow(b, graphs=True)
    HS = SqOp.from_numpy(b, H, 1)
    G = SqOp.zeros(b, -1)
    f.magnus_decouple(HS, G)
    Op = SqOp.from_numpy(b, H, 1)
    f.transform_sq_op_BCH(Op, G)
    torch.cuda.synchronize()
    assert_rel(Op.numpy(), HS.numpy(), rtol=1e-12, what="transformed bare H")
    # an observable (another hermitian operator) under the flow's Omega
    Obs = _hamiltonian(emax, seed=9, hw=16.0, scale=0.5)
    Gn = G.numpy()
    ref, it, _ = o.bch(Gn, Obs, REF)
    O = SqOp.from_numpy(b, Obs, 1)
    nt = f.transform_sq_op_BCH(O, G)
    torch.cuda.synchronize()
    assert nt == it
    assert_rel(O.numpy(), ref, what="transformed observable")


@pytest.mark.parametrize("gen", ["white", "w2"])
def test_dhds(gpu, gen):
    """dHds: DH = [eta(H), H] against the oracle's generator and commutator"""
    from eom_imsrg import Flow, SqOp
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    print(f"[dHds] emax=4 generator={gen}", flush=True)
    H = _hamiltonian(emax)
    eta = o.white(H) if gen == "white" else o.gs_w2(H)
    ref = o.commutator(eta, -1, H, 1, 1)
    f = Flow(b, generator=gen)
    DH = SqOp.zeros(b)
    Hd = SqOp.from_numpy(b, H, 1)
    f.dHds(Hd, DH)
    torch.cuda.synchronize()
    assert_rel(DH.numpy(), ref, what=f"dHds {gen}")
    assert np.array_equal(Hd.numpy(), H)
