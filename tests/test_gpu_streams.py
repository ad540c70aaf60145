"""The concurrent term schedule of the grouped commutator (GEMM families,
the Pandya transform of R and the zero- and one-body terms on auxiliary
streams) gives bit-identical results to the single-stream schedule: every
output element is accumulated in the same order."""
import numpy as np
import pytest
import torch

from conftest import basis

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("emax", [4, 6])
@pytest.mark.parametrize("graphs", [True, False])
def test_bch_streams_bitwise(gpu, emax, graphs):
    """BCH_EXPAND (a full commutator per term): schedules on, off, on"""
    from eom_imsrg import Flow, SqOp, commutators as cm, synth
    b = basis(emax)
    print(f"[streams] bch emax={emax} graphs={graphs}", flush=True)
    G = SqOp.from_numpy(b, synth.antihermitian(b, seed=3, sigma=0.05), -1)
    H = SqOp.from_numpy(b, synth.hamiltonian_vacuum(b, seed=1, hw=20.0), 1)
    cm.normal_order(H)
    out = []
    try:
        for on in (True, False, True):
            b.set_term_streams(on)
            HS = SqOp.zeros(b, 1)
            n, adv = Flow(b, graphs=graphs).BCH_EXPAND(HS, G, H)
            torch.cuda.synchronize()
            out.append((n, HS.numpy(), adv))
    finally:
        b.set_term_streams(True)
    assert out[0][0] == out[1][0] == out[2][0] and out[0][0] > 3
    assert np.array_equal(out[0][2], out[1][2]) and np.array_equal(out[0][2], out[2][2])
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][1], out[2][1])


@pytest.mark.parametrize("graphs", [True, False])
def test_flow_streams_bitwise(gpu, graphs):
    """the whole flow (BCH and Magnus terms, captured in graphs or not)"""
    from eom_imsrg import Flow, SqOp, commutators as cm, synth
    b = basis(4)
    print(f"[streams] flow graphs={graphs}", flush=True)
    H0 = synth.hamiltonian_vacuum(b, seed=0, hw=20.0)
    logs, HS = [], []
    try:
        for on in (True, False):
            b.set_term_streams(on)
            H = SqOp.from_numpy(b, H0, 1)
            cm.normal_order(H)
            logs.append(Flow(b, graphs=graphs).magnus_decouple(H, SqOp.zeros(b, -1)))
            torch.cuda.synchronize()
            HS.append(H.numpy())
    finally:
        b.set_term_streams(True)
    assert np.array_equal(logs[0], logs[1])
    assert np.array_equal(HS[0], HS[1])


def test_flow_gradient_streams(gpu):
    """the reverse sweep with the 222_ph chain and the dense 122 cotangent of
    each commutator transpose on auxiliary streams equals the single-stream
    sweep to rounding (the 222_ph scatter transpose accumulates its 64 x 64
    W' cotangent tiles with LDS atomics, so the sweep is reproducible only to
    rounding in either schedule; a missing wait would show as O(1) errors)"""
    from eom_imsrg import Flow, adjoint, synth
    b = basis(4)
    print("[streams] flow gradient", flush=True)
    T = torch.from_numpy(synth.hamiltonian_vacuum(b, 0)).cuda()
    out = []
    try:
        for on in (True, False):
            b.set_term_streams(on)
            e0, Tb, st = adjoint.flow_grad(Flow(b, graphs=True), T, ckpt=4)
            torch.cuda.synchronize()
            out.append((e0, st, Tb.cpu().numpy()))
    finally:
        b.set_term_streams(True)
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    from conftest import assert_rel
    assert_rel(out[0][2], out[1][2], rtol=1e-11, what="dE0/dT streams on vs off")
