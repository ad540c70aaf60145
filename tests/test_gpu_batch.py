"""Many-instance batches on the GPU (imsrg_batch_*): every row of a batched
call against a single-instance call on the same inputs (the kernels are the
same, so the rows must agree to the last bits, and the data-dependent exits
exactly), and the batched flow's E0 against the oracle's flow.

Inputs follow the batch recipe: instance k has seed k, two-body scale
U(0.8, 1.2) and hw in {16, 20, 24} (synth.instance_parameters).
"""
import numpy as np
import pytest
import torch

from conftest import assert_rel, basis, oracle_basis

pytestmark = pytest.mark.gpu

REF = float(np.float32(1e-6))


def _instances(emax, ninst):
    from eom_imsrg import synth
    b, o = basis(emax), oracle_basis(emax)
    Hs = []
    for k in range(ninst):
        seed, scale, hw = synth.instance_parameters(k)
        Hs.append(o.normal_order(synth.hamiltonian_vacuum(b, seed=seed, hw=hw, v_scale=scale)))
    return np.stack(Hs)


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()


def test_batch_commutator_rows_match_single(gpu):
    from eom_imsrg import Batch, SqOp, commutators as cm, synth
    emax, ni = 4, 3
    b = basis(emax)
    print(f"[batch commutator] emax={emax} ninst={ni}", flush=True)
    L = np.stack([synth.antihermitian(b, 10 + k, 0.1) for k in range(ni)])
    R = _instances(emax, ni)
    bt = Batch(b, ni, graphs=False)
    RES = bt.zeros()
    bt.commutator(_t(L), -1, _t(R), 1, RES)
    torch.cuda.synchronize()
    for k in range(ni):
        one = SqOp.zeros(b, 1)
        cm.commutator(SqOp.from_numpy(b, L[k], -1), SqOp.from_numpy(b, R[k], 1), one)
        torch.cuda.synchronize()
        assert_rel(RES[k].cpu().numpy(), one.numpy(), rtol=1e-14, what=f"row {k}")


@pytest.mark.parametrize("graphs", [False, True], ids=["eager", "graph"])
def test_batch_bch_and_magnus_rows_match_single(gpu, graphs):
    from eom_imsrg import Batch, Flow, SqOp, synth
    emax, ni = 4, 3
    b, o = basis(emax), oracle_basis(emax)
    print(f"[batch bch/magnus] graphs={graphs}", flush=True)
    H = _instances(emax, ni)
    # different generator sizes give different term counts per instance
    G = np.stack([synth.antihermitian(b, 20 + k, 0.02 * (k + 1)) for k in range(ni)])
    bt = Batch(b, ni, graphs=graphs)
    HS = bt.zeros()
    nt = bt.BCH_EXPAND(HS, _t(G), _t(H), conv=1e-10)
    eta = np.stack([o.white(H[k]) for k in range(ni)])
    DG = _t(eta)
    nc = bt.MAGNUS_EXPAND(DG, _t(G), conv=1e-10)
    torch.cuda.synchronize()
    f = Flow(b, graphs=graphs)
    for k in range(ni):
        hs = SqOp.zeros(b)
        it, _ = f.BCH_EXPAND(hs, SqOp.from_numpy(b, G[k], -1), SqOp.from_numpy(b, H[k], 1), conv=1e-10)
        dg = SqOp.from_numpy(b, eta[k], -1)
        ncs = f.MAGNUS_EXPAND(dg, SqOp.from_numpy(b, G[k], -1), conv=1e-10)
        torch.cuda.synchronize()
        assert nt[k] == it and nc[k] == ncs, (k, nt, it, nc, ncs)
        assert_rel(HS[k].cpu().numpy(), hs.numpy(), rtol=1e-14, what=f"BCH row {k}")
        assert_rel(DG[k].cpu().numpy(), dg.numpy(), rtol=1e-14, what=f"Magnus row {k}")
    assert len(set(nt.tolist())) > 1, f"term counts {nt} do not exercise the per-instance exits"


def test_batch_magnus_decouple_emax4(gpu):
    """4 instances flowed together: per-instance E0 and step counts equal the
    single-instance flows; two instances against the oracle's flow (1e-8 MeV)"""
    from eom_imsrg import Batch, Flow, SqOp
    emax, ni = 4, 4
    b, o = basis(emax), oracle_basis(emax)
    print(f"[batch decouple] emax={emax} ninst={ni}", flush=True)
    H = _instances(emax, ni)
    bt = Batch(b, ni, graphs=True)
    HS = _t(H)
    G = bt.zeros()
    e0, steps, nmax = bt.magnus_decouple(HS, G)
    torch.cuda.synchronize()
    print("  E0", e0, "steps", steps, "batched steps", nmax, flush=True)
    assert nmax == steps.max()
    f = Flow(b, graphs=True)
    for k in range(ni):
        hs = SqOp.from_numpy(b, H[k], 1)
        g = SqOp.zeros(b, -1)
        log = f.magnus_decouple(hs, g)
        torch.cuda.synchronize()
        assert len(log) == steps[k], (k, len(log), steps[k])
        assert abs(log[-1, 1] - e0[k]) <= 1e-12 * abs(e0[k]), (k, log[-1, 1], e0[k])
        assert_rel(HS[k].cpu().numpy(), hs.numpy(), rtol=1e-12, what=f"flowed H row {k}")
    for k in (0, ni - 1):
        HSr, _, logr, str_ = o.magnus_decouple(H[k], 0.5, REF, REF, 0, REF)
        assert str_ == steps[k]
        assert abs(HSr[0] - e0[k]) < 1e-8, (k, HSr[0], e0[k])


def test_batch_rejects_bad_use(gpu):
    from eom_imsrg import Batch, Flow, SqOp, _lib
    b = basis(2)
    bt = Batch(b, 2)
    with pytest.raises(ValueError):
        bt.commutator(bt.zeros()[:1], -1, bt.zeros(), 1, bt.zeros())
    # a batched flow refuses the single-instance drivers and vice versa
    hs = SqOp.zeros(b)
    rc = _lib.lib.imsrg_bch_expand(bt.handle, hs.ptr, hs.ptr, -1, hs.ptr, 1, 1e-6, 40, None, None, None)
    assert rc == 1
    f = Flow(b)
    rc = _lib.lib.imsrg_batch_magnus_expand(f.handle, hs.ptr, hs.ptr, 1e-6, None, None)
    assert rc == 1
