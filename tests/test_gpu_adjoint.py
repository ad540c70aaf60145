"""Reverse mode on the GPU: the hand-written transposes (imsrg_*_vjp,
imsrg_flow_grad) against the oracle's exact transposes on the same inputs,
against directional derivatives of the GPU forward commutator (bilinear, so
exact), and the flow gradient against central differences of the GPU flow.

Gradients are compared after folding onto the independent (anti)hermitian
matrix elements (adjoint.symmetric_gradient): the flow only ever perturbs
operators along those directions.  Tolerance 1e-9 relative per element for
single maps (atomic adds change the summation order), 1e-8 for the flow.
"""
import numpy as np
import pytest
import torch

from conftest import assert_rel, basis, oracle_basis

pytestmark = pytest.mark.gpu

HERMS = [(-1, 1), (-1, -1), (1, 1), (1, -1)]
REF = float(np.float32(1e-6))


def _rand(b, h, seed, sigma=1.0):
    from eom_imsrg import synth
    return synth.hermitian_random(b, seed, sigma) if h == 1 else synth.antihermitian(b, seed, sigma)


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()


def _sym(b, g, h):
    from eom_imsrg import adjoint
    return adjoint.symmetric_gradient(b, np.asarray(g), h)


@pytest.mark.parametrize("emax", [2, 4])
@pytest.mark.parametrize("hL,hR", HERMS)
def test_commutator_vjp(gpu, emax, hL, hR):
    from eom_imsrg import SqOp, adjoint, commutators as cm
    b, o = basis(emax), oracle_basis(emax)
    print(f"[commutator_vjp] emax={emax} herm=({hL},{hR})", flush=True)
    L, R = _rand(b, hL, 3), _rand(b, hR, 4)
    hres = -hL * hR
    x = np.random.default_rng(5).normal(size=b.oplen)
    Lb_o, Rb_o = o.commutator_vjp(L, hL, R, hR, x, hres)
    Ld, Rd = SqOp.from_numpy(b, L, hL), SqOp.from_numpy(b, R, hR)
    Lb, Rb = torch.zeros(b.oplen, dtype=torch.float64, device="cuda"), torch.zeros(b.oplen, dtype=torch.float64,
                                                                                     device="cuda")
    adjoint.commutator_vjp(Ld, Rd, _t(x), hres, Lb, Rb)
    torch.cuda.synchronize()
    lb, rb = Lb.cpu().numpy(), Rb.cpu().numpy()
    assert_rel(_sym(b, lb, hL), _sym(b, Lb_o, hL), rtol=1e-9, what="Lbar")
    assert_rel(_sym(b, rb, hR), _sym(b, Rb_o, hR), rtol=1e-9, what="Rbar")
    # directional derivatives of the GPU forward commutator
    for s in range(2):
        dL, dR = _rand(b, hL, 70 + s), _rand(b, hR, 80 + s)
        RES = SqOp.zeros(b, hres)
        cm.commutator(SqOp.from_numpy(b, dL, hL), Rd, RES)
        r1 = x @ RES.numpy()
        cm.commutator(Ld, SqOp.from_numpy(b, dR, hR), RES)
        r2 = x @ RES.numpy()
        assert abs(lb @ dL - r1) <= 1e-10 * (abs(r1) + np.abs(x).sum() * 1e-3), (lb @ dL, r1)
        assert abs(rb @ dR - r2) <= 1e-10 * (abs(r2) + np.abs(x).sum() * 1e-3), (rb @ dR, r2)


@pytest.mark.parametrize("tile", [0, 5, 6, 7, 6 | 8])
def test_commutator_vjp_tile_shapes(gpu, tile):
    """the transposed 222_pp_hh products (the pp-output problems and the
    rounds of the others) and the W' transpose on every tile shape
    (LDS-staged and register-direct, one kernel per transpose pair):
    equal to the oracle's transposes at eMax 4 and to the default tile
    shape's result at eMax 6 (blocks larger than the tiles)"""
    from eom_imsrg import SqOp, _lib, adjoint
    hL, hR, hres = -1, 1, 1
    for emax in (4, 6):
        b = basis(emax)
        print(f"[commutator_vjp tiles] emax={emax} tile={tile}", flush=True)
        L, R = _rand(b, hL, 3), _rand(b, hR, 4)
        x = np.random.default_rng(5).normal(size=b.oplen)
        Ld, Rd = SqOp.from_numpy(b, L, hL), SqOp.from_numpy(b, R, hR)
        out = []
        for t in (3, tile):
            for fam in ("ADJ", "ADJ_PH", "ADJ_PP"):
                _lib.check(_lib.lib.imsrg_set_gemm_tile(b.handle, _lib.GEMM[fam], t))
            Lb = torch.zeros(b.oplen, dtype=torch.float64, device="cuda")
            Rb = torch.zeros_like(Lb)
            adjoint.commutator_vjp(Ld, Rd, _t(x), hres, Lb, Rb)
            torch.cuda.synchronize()
            out.append((Lb.cpu().numpy(), Rb.cpu().numpy()))
        for fam, t in (("ADJ", 5), ("ADJ_PH", 6), ("ADJ_PP", 11)):
            _lib.check(_lib.lib.imsrg_set_gemm_tile(b.handle, _lib.GEMM[fam], t))
        (l0, r0), (l1, r1) = out
        assert_rel(_sym(b, l1, hL), _sym(b, l0, hL), rtol=1e-11, what=f"Lbar tile {tile} vs 3")
        assert_rel(_sym(b, r1, hR), _sym(b, r0, hR), rtol=1e-11, what=f"Rbar tile {tile} vs 3")
        if emax == 4:
            Lb_o, Rb_o = oracle_basis(emax).commutator_vjp(L, hL, R, hR, x, hres)
            assert_rel(_sym(b, l1, hL), _sym(b, Lb_o, hL), rtol=1e-9, what="Lbar vs oracle")
            assert_rel(_sym(b, r1, hR), _sym(b, Rb_o, hR), rtol=1e-9, what="Rbar vs oracle")


def test_bch_magnus_white_normal_order_vjp(gpu):
    from eom_imsrg import Flow, SqOp, adjoint, synth
    emax = 4
    b, o = basis(emax), oracle_basis(emax)
    print("[flow maps vjp] emax=4", flush=True)
    T = synth.hamiltonian_vacuum(b, 0)
    H = o.normal_order(T)
    G = synth.antihermitian(b, 3, 0.05)
    rng = np.random.default_rng(11)
    f = Flow(b)
    zeros = lambda: torch.zeros(b.oplen, dtype=torch.float64, device="cuda")
    # BCH
    HSb = rng.normal(size=b.oplen)
    Gb_o, Hb_o, it_o = o.bch_vjp(G, H, 1e-10, HSb)
    Gb, Hb = zeros(), zeros()
    it = adjoint.bch_vjp(f, SqOp.from_numpy(b, G, -1), SqOp.from_numpy(b, H, 1), _t(HSb), Gb, Hb, conv=1e-10)
    torch.cuda.synchronize()
    assert it == it_o
    assert_rel(_sym(b, Gb.cpu().numpy(), -1), _sym(b, Gb_o, -1), rtol=1e-9, what="BCH Gbar")
    assert_rel(_sym(b, Hb.cpu().numpy(), 1), _sym(b, Hb_o, 1), rtol=1e-9, what="BCH Hbar")
    # Magnus
    eta = o.white(H)
    G2 = synth.antihermitian(b, 6, 0.3)
    DGb = rng.normal(size=b.oplen)
    Eb_o, Gb2_o, k_o = o.magnus_vjp(eta, G2, 1e-12, DGb)
    Eb, Gb2 = zeros(), zeros()
    k = adjoint.magnus_vjp(f, SqOp.from_numpy(b, eta, -1), SqOp.from_numpy(b, G2, -1), _t(DGb), Eb, Gb2, conv=1e-12)
    torch.cuda.synchronize()
    assert k == k_o
    assert_rel(_sym(b, Eb.cpu().numpy(), -1), _sym(b, Eb_o, -1), rtol=1e-9, what="Magnus etabar")
    assert_rel(_sym(b, Gb2.cpu().numpy(), -1), _sym(b, Gb2_o, -1), rtol=1e-9, what="Magnus Gbar")
    # white
    ETAb = rng.normal(size=b.oplen)
    Hb3 = zeros()
    adjoint.white_vjp(SqOp.from_numpy(b, H, 1), _t(ETAb), Hb3)
    torch.cuda.synchronize()
    assert_rel(_sym(b, Hb3.cpu().numpy(), 1), _sym(b, o.white_vjp(H, ETAb), 1), rtol=1e-9, what="white Hbar")
    # normal ordering
    Hb4 = rng.normal(size=b.oplen)
    Tb = zeros()
    adjoint.normal_order_vjp(b, _t(Hb4), Tb)
    torch.cuda.synchronize()
    assert_rel(_sym(b, Tb.cpu().numpy(), 1), _sym(b, o.normal_order_vjp(Hb4), 1), rtol=1e-12, what="normal order")


@pytest.mark.parametrize("emax", [2, 4])
def test_flow_grad_vs_oracle(gpu, emax):
    """the GPU reverse sweep with Omega every 4 steps (segments recomputed),
    every 3 steps, and every step's states kept (ckpt 0, nothing recomputed)
    against the oracle's reverse sweep"""
    from eom_imsrg import Flow, adjoint, synth
    b, o = basis(emax), oracle_basis(emax)
    T = synth.hamiltonian_vacuum(b, 0)
    e0_o, Tb_o, st_o = o.flow_grad(T, 0.5, REF, REF, REF, ckpt=4)
    f = Flow(b, graphs=True)
    for ck in (4, 3, 0):
        print(f"[flow_grad] emax={emax} ckpt={ck}", flush=True)
        e0, Tb, st = adjoint.flow_grad(f, _t(T), ckpt=ck)
        assert st == st_o and abs(e0 - e0_o) < 1e-8
        assert_rel(_sym(b, Tb.cpu().numpy(), 1), _sym(b, Tb_o, 1), rtol=1e-8, what=f"dE0/dT (ckpt {ck})")


def test_flow_energy_autograd_and_central_differences(gpu):
    """torch.autograd.Function over the flow; central differences (h = 1e-5)
    of the GPU flow on hermitian two-body matrix elements, eMax 4"""
    from eom_imsrg import Flow, flow_energy, synth
    from jscheme import each_element
    b = basis(4)
    print("[flow_energy] emax=4", flush=True)
    f = Flow(b, graphs=True)
    T0 = _t(synth.hamiltonian_vacuum(b, 0))
    T = T0.clone().requires_grad_(True)
    E = flow_energy(T, f, ckpt=4)
    E.backward()
    g = T.grad.cpu().numpy()
    gs = _sym(b, g, 1)
    h = 1e-5
    rng = np.random.default_rng(3)
    elems = [e for e in each_element(b) if e[1] in (0, 2, 3, 4)]
    big = [e for e in elems if abs(gs[int(b.goff[e[0], e[1]]) + e[2] + e[3] * b.block_shape(e[0], e[1])[0]]) > 1e-3]
    for k in rng.choice(len(big), 4, replace=False):
        q, gg, ix, jx, _, _, _ = big[k]
        n1, _ = b.block_shape(q, gg)
        d = torch.zeros(b.oplen, dtype=torch.float64, device="cuda")
        d[int(b.goff[q, gg]) + ix + jx * n1] = 1.0
        if gg != 2:
            d[int(b.goff[q, gg]) + jx + ix * n1] = 1.0
        with torch.no_grad():
            ep = flow_energy(T0 + h * d, f).item()
            em = flow_energy(T0 - h * d, f).item()
        fd = (ep - em) / (2 * h)
        an = float(g @ d.cpu().numpy())
        assert abs(fd - an) < 1e-6 * abs(fd) + 5e-8, (fd, an)


def test_flow_grad_emax8_vs_cpu_reference_fixture(gpu):
    """dE0/dT of the whole eMax 8 flow (Omega checkpointed every 4 steps)
    against the oracle's reverse sweep run on the CPU
    (tests/golden/grad_emax8.json, written by make_grad_emax8.py): the 200
    largest and 2000 seeded further elements of the folded gradient, 1e-6
    relative; E0 and the step count of the forward sweep exactly as there."""
    import json
    import os
    from eom_imsrg import Flow, adjoint, synth
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "grad_emax8.json")))
    b = basis(8)
    print("[flow_grad] emax=8 vs CPU fixture", flush=True)
    T = synth.hamiltonian_vacuum(b, seed=0, hw=20.0)
    e0, Tb, st = adjoint.flow_grad(Flow(b, graphs=True), _t(T), ckpt=4)
    assert st == fx["steps"] and abs(e0 - fx["E0"]) < 1e-8, (st, e0, fx["E0"])
    g = _sym(b, Tb.cpu().numpy(), 1)
    idx = np.asarray(fx["index"], np.int64)
    ref = np.asarray(fx["grad"])
    assert_rel(g[idx], ref, rtol=1e-6, what="dE0/dT vs CPU reference gradient")
    assert abs(np.linalg.norm(g) - fx["norm"]) <= 1e-6 * fx["norm"]
