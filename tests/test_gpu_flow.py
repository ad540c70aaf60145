"""BCH / Magnus / full-flow drivers on the GPU vs the CPU oracle.

Flowed E0 must agree within 1e-8 MeV (the configurations' bar); the
operators within 1e-10 relative per element; the data-dependent term
counts must agree exactly.
"""
import numpy as np
import pytest
import torch

from conftest import assert_rel, basis, oracle_basis

pytestmark = pytest.mark.gpu

REF = float(np.float32(1e-6))


def _hamiltonian(emax, seed=0, hw=20.0):
    from eom_imsrg import synth
    b, o = basis(emax), oracle_basis(emax)
    return o.normal_order(synth.hamiltonian_vacuum(b, seed=seed, hw=hw))


@pytest.mark.parametrize("graphs", [False, True], ids=["eager", "graph"])
@pytest.mark.parametrize("path", [0, 1], ids=["grouped", "plain"])
def test_bch_expand(gpu, graphs, path):
    from eom_imsrg import Flow, SqOp, synth
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    if path == 1 and graphs:
        pytest.skip("graphs apply to the grouped path")
    b.set_path(path)
    print(f"[bch] graphs={graphs} path={path}", flush=True)
    H = _hamiltonian(emax)
    G = synth.antihermitian(b, 3, 0.05)
    HSr, itr, advr = o.bch(G, H, 1e-10)
    f = Flow(b, graphs=graphs)
    HS = SqOp.zeros(b)
    it, adv = f.BCH_EXPAND(HS, SqOp.from_numpy(b, G, -1), SqOp.from_numpy(b, H, 1), conv=1e-10)
    torch.cuda.synchronize()
    assert it == itr
    assert_rel(adv, advr[: it - 1], rtol=1e-9, what="BCH term norms")
    assert_rel(HS.numpy(), HSr, what="BCH_EXPAND")
    # a second call reuses any captured graphs and must give the same bits
    HS2 = SqOp.zeros(b)
    f.BCH_EXPAND(HS2, SqOp.from_numpy(b, G, -1), SqOp.from_numpy(b, H, 1), conv=1e-10)
    torch.cuda.synchronize()
    assert torch.equal(HS.data, HS2.data)
    b.set_path(0)


def test_magnus_expand(gpu):
    from eom_imsrg import Flow, SqOp, synth
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    H = _hamiltonian(emax)
    eta = o.white(H)
    G = synth.antihermitian(b, 4, 0.05)
    ref, nref = o.magnus(eta, G, REF)
    f = Flow(b)
    DG = SqOp.from_numpy(b, eta, -1)
    n = f.MAGNUS_EXPAND(DG, SqOp.from_numpy(b, G, -1))
    torch.cuda.synchronize()
    assert n == nref
    assert_rel(DG.numpy()[1:], ref[1:], what="MAGNUS_EXPAND")
    # Omega = 0: dOmega = eta exactly (IMSRG_MAGNUS.f90:654-655)
    DG0 = SqOp.from_numpy(b, eta, -1)
    assert f.MAGNUS_EXPAND(DG0, SqOp.zeros(b, -1)) == 0
    assert np.array_equal(DG0.numpy(), eta)


@pytest.mark.parametrize("graphs", [False, True], ids=["eager", "graph"])
def test_magnus_decouple_emax4(gpu, graphs):
    """config 1's flow (eMax 4, 16O holes, hw 20, seed 0) to |mbpt2| < 1e-6"""
    from eom_imsrg import Flow, SqOp
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    print(f"[flow] emax=4 graphs={graphs}", flush=True)
    H = _hamiltonian(emax)
    HSr, Gr, logr, str_ = o.magnus_decouple(H, 0.5, REF, REF, 0, REF)
    f = Flow(b, graphs=graphs)
    HS = SqOp.from_numpy(b, H, 1)
    G = SqOp.zeros(b, -1)
    log = f.magnus_decouple(HS, G, ds=0.5)
    torch.cuda.synchronize()
    assert len(log) == str_
    assert np.array_equal(log[:, 3], logr[:, 3])          # nested commutators per step
    assert abs(HS.E0 - HSr[0]) < 1e-8, (HS.E0, HSr[0])
    assert np.max(np.abs(log[:, 1] - logr[:, 1])) < 1e-8
    assert_rel(HS.numpy(), HSr, rtol=1e-9, what="decoupled H")
    assert_rel(G.numpy(), Gr, rtol=1e-9, what="Omega")


def test_flow_matches_committed_fixture(gpu):
    """the eMax-4 flow from the product's own normal ordering against the
    committed fixture (tests/golden/oracle_emax4.json): same step count and
    commutators per step, E0 per step within 1e-8 MeV"""
    import json
    import os
    from eom_imsrg import Flow, SqOp, commutators as cm, synth
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_emax4.json")))
    b = basis(4)
    print("[fixture flow] emax=4", flush=True)
    H = SqOp.from_numpy(b, synth.hamiltonian_vacuum(b, seed=0, hw=20.0), 1)
    cm.normal_order(H)
    torch.cuda.synchronize()
    assert abs(H.E0 - g["E0_bare_normal_ordered"]) < 1e-9
    assert abs(cm.mbpt2(H) - g["mbpt2_bare"]) < 1e-9
    f = Flow(b, graphs=True)
    G = SqOp.zeros(b, -1)
    log = f.magnus_decouple(H, G, ds=0.5)
    torch.cuda.synchronize()
    assert len(log) == g["flow"]["steps"]
    assert [int(x) for x in log[:, 3]] == g["flow"]["commutators"]
    assert np.abs(log[:, 1] - np.array(g["flow"]["E0"])).max() < 1e-8
    assert abs(H.E0 - g["flow"]["final_E0"]) < 1e-8
