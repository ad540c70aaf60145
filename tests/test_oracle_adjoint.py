"""Pins of the oracle's reverse-mode derivatives (oracle/imsrg_adjoint.inc),
the CPU reference gradient of the flow:

* commutator: bilinear, so <RESbar, [e_x, R]> must equal Lbar[x] (and the
  same for R) for EVERY flat unit vector e_x, all four herm pairs, eMax 2;
* normal_order: linear, checked on random directions;
* BCH, Magnus, white: central differences on random directions;
* the flow gradient dE0/dT: central differences (h = 1e-5 MeV) on random
  two-body directions and on single hermitian matrix elements, eMax 2 and 4.
"""
import numpy as np
import pytest

from conftest import basis, oracle_basis
from jscheme import each_element, two_body_part

HERMS = [(-1, 1), (-1, -1), (1, 1), (1, -1)]
REF = float(np.float32(1e-6))


def _rand(b, h, seed, sigma=1.0):
    from eom_imsrg import synth
    return synth.hermitian_random(b, seed, sigma) if h == 1 else synth.antihermitian(b, seed, sigma)


@pytest.mark.parametrize("hL,hR", HERMS)
def test_commutator_vjp_every_unit_vector(hL, hR):
    b, o = basis(2), oracle_basis(2)
    L, R = _rand(b, hL, 3), _rand(b, hR, 4)
    hres = -hL * hR
    RESb = np.random.default_rng(5).normal(size=b.oplen)
    Lb, Rb = o.commutator_vjp(L, hL, R, hR, RESb, hres)
    n = b.oplen
    gl, gr = np.zeros(n), np.zeros(n)
    e = np.zeros(n)
    for x in range(n):
        e[x] = 1.0
        gl[x] = RESb @ o.commutator(e, hL, R, hR, hres)
        gr[x] = RESb @ o.commutator(L, hL, e, hR, hres)
        e[x] = 0.0
    scale = max(np.abs(gl).max(), np.abs(gr).max())
    assert np.abs(Lb - gl).max() < 1e-11 * scale, np.abs(Lb - gl).max()
    assert np.abs(Rb - gr).max() < 1e-11 * scale, np.abs(Rb - gr).max()
    assert np.abs(Lb).max() > 1e-3 and np.abs(Rb).max() > 1e-3


def test_normal_order_vjp_linear():
    from eom_imsrg import synth
    b, o = basis(4), oracle_basis(4)
    T = synth.hamiltonian_vacuum(b, 0)
    rng = np.random.default_rng(2)
    Hb = rng.normal(size=b.oplen)
    Tb = o.normal_order_vjp(Hb)
    for _ in range(4):
        d = rng.normal(size=b.oplen)
        lhs = Hb @ (o.normal_order(T + d) - o.normal_order(T))
        assert abs(lhs - Tb @ d) < 1e-10 * (abs(lhs) + 1)


def _central(f, x, d, h):
    return (f(x + h * d) - f(x - h * d)) / (2 * h)


def test_bch_vjp():
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    G = synth.antihermitian(b, 3, 0.05)
    HSb = np.random.default_rng(1).normal(size=b.oplen)
    Gb, Hb, it = o.bch_vjp(G, H, 1e-10, HSb)
    assert it == o.bch(G, H, 1e-10)[1]
    for s in range(3):
        dG = synth.antihermitian(b, 10 + s, 0.05)
        dH = synth.hermitian_random(b, 20 + s, 1.0)
        fd_g = _central(lambda g: HSb @ o.bch(g, H, 1e-10)[0], G, dG, 1e-5)
        fd_h = _central(lambda hh: HSb @ o.bch(G, hh, 1e-10)[0], H, dH, 1e-5)
        assert abs(fd_g - Gb @ dG) < 1e-7 * (abs(fd_g) + 1e-3), (fd_g, Gb @ dG)
        assert abs(fd_h - Hb @ dH) < 1e-7 * (abs(fd_h) + 1e-3), (fd_h, Hb @ dH)


def test_magnus_vjp():
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    eta = o.white(H)
    G = synth.antihermitian(b, 6, 0.3)
    DGb = np.random.default_rng(7).normal(size=b.oplen)
    Eb, Gb, k = o.magnus_vjp(eta, G, 1e-12, DGb)
    assert k >= 3
    for s in range(3):
        dG = synth.antihermitian(b, 30 + s, 0.05)
        dE = synth.antihermitian(b, 40 + s, 0.05)
        fd_g = _central(lambda g: DGb @ o.magnus(eta, g, 1e-12)[0], G, dG, 1e-5)
        fd_e = _central(lambda e: DGb @ o.magnus(e, G, 1e-12)[0], eta, dE, 1e-5)
        assert abs(fd_g - Gb @ dG) < 1e-7 * (abs(fd_g) + 1e-3), (fd_g, Gb @ dG)
        assert abs(fd_e - Eb @ dE) < 1e-7 * (abs(fd_e) + 1e-3), (fd_e, Eb @ dE)


def test_white_vjp():
    from eom_imsrg import synth
    b, o = basis(4), oracle_basis(4)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    Eb = np.random.default_rng(3).normal(size=b.oplen)
    Hb = o.white_vjp(H, Eb)
    for s in range(3):
        dH = synth.hermitian_random(b, 50 + s, 0.1)
        fd = _central(lambda h: Eb @ o.white(h), H, dH, 1e-5)
        assert abs(fd - Hb @ dH) < 1e-7 * (abs(fd) + 1e-3), (fd, Hb @ dH)


def _two_body_direction(b, seed):
    """hermitian perturbation of the vacuum two-body part only"""
    from eom_imsrg import synth
    d = synth.hermitian_random(b, seed, 1.0)
    d[: two_body_part(b).start] = 0.0
    return d


def _flow_e0(o, T):
    H = o.normal_order(T)
    HS, _, log, st = o.magnus_decouple(H, 0.5, REF, REF, 0, REF)
    return HS[0], st


@pytest.mark.parametrize("emax,ndir", [(2, 4), (4, 2)])
def test_flow_gradient_vs_central_differences(emax, ndir):
    from eom_imsrg import synth
    b, o = basis(emax), oracle_basis(emax)
    T = synth.hamiltonian_vacuum(b, 0)
    e0, Tb, st = o.flow_grad(T, 0.5, REF, REF, REF, ckpt=4)
    e0_ref, st_ref = _flow_e0(o, T)
    assert st == st_ref and abs(e0 - e0_ref) < 1e-10
    h = 1e-5
    for s in range(ndir):
        d = _two_body_direction(b, 100 + s)
        ep, sp = _flow_e0(o, T + h * d)
        em, sm = _flow_e0(o, T - h * d)
        assert sp == sm == st
        fd = (ep - em) / (2 * h)
        assert abs(fd - Tb @ d) < 1e-6 * abs(fd), (fd, Tb @ d)
    # single elements (with their hermitian mirror for square blocks); E0 is
    # a few hundred MeV, so the difference quotient carries ~1e-8 of rounding
    rng = np.random.default_rng(9)
    elems = [t for t in each_element(b) if t[1] in (0, 2, 3, 4)
             and abs(Tb[int(b.goff[t[0], t[1]]) + t[2] + t[3] * b.block_shape(t[0], t[1])[0]]) > 1e-3]
    for k in rng.choice(len(elems), 3, replace=False):
        q, g, ix, jx, _, _, _ = elems[k]
        n1, _ = b.block_shape(q, g)
        d = np.zeros(b.oplen)
        d[int(b.goff[q, g]) + ix + jx * n1] = 1.0
        if g != 2:
            d[int(b.goff[q, g]) + jx + ix * n1] = 1.0
        ep, _ = _flow_e0(o, T + h * d)
        em, _ = _flow_e0(o, T - h * d)
        fd = (ep - em) / (2 * h)
        assert abs(fd - Tb @ d) < 1e-6 * abs(fd) + 5e-8, (fd, Tb @ d)


def test_atan_atanxx_vjp():
    """gen_core transpose in atan mode against central differences: aliased
    (build_gs_atan, Hd = Hn) and with separate directions for the
    denominator operator and the numerator operator (build_gs_atanxx)"""
    from eom_imsrg import synth
    b, o = basis(4), oracle_basis(4)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    H2 = o.commutator(synth.antihermitian(b, 7, 0.1), -1, H, 1, 1)
    Eb = np.random.default_rng(4).normal(size=b.oplen)
    Hdb, Hnb = o.gen_vjp(H, H, Eb, 1)
    for s in range(2):
        dH = synth.hermitian_random(b, 60 + s, 0.1)
        fd = _central(lambda h: Eb @ o.gs_atan(h), H, dH, 1e-5)
        assert abs(fd - (Hdb + Hnb) @ dH) < 1e-7 * (abs(fd) + 1e-3), (fd, (Hdb + Hnb) @ dH)
    Hdb, Hnb = o.gen_vjp(H, H2, Eb, 1)
    for s in range(2):
        dH = synth.hermitian_random(b, 70 + s, 0.1)
        fd = _central(lambda h: Eb @ o.gs_atanxx(h, H2), H, dH, 1e-5)
        assert abs(fd - Hdb @ dH) < 1e-7 * (abs(fd) + 1e-3), ("Hold", fd, Hdb @ dH)
        fd = _central(lambda h: Eb @ o.gs_atanxx(H, h), H2, dH, 1e-5)
        assert abs(fd - Hnb @ dH) < 1e-7 * (abs(fd) + 1e-3), ("H", fd, Hnb @ dH)
    # mode 0 with Hd = Hn is the white transpose
    Hdb, Hnb = o.gen_vjp(H, H, Eb, 0)
    assert np.allclose(Hdb + Hnb, o.white_vjp(H, Eb), rtol=1e-13, atol=1e-15)


def test_w2_vjp():
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    Eb = np.random.default_rng(5).normal(size=b.oplen)
    Hb = o.gs_w2_vjp(H, Eb)
    for s in range(3):
        dH = synth.hermitian_random(b, 80 + s, 0.1)
        fd = _central(lambda h: Eb @ o.gs_w2(h), H, dH, 1e-5)
        assert abs(fd - Hb @ dH) < 1e-7 * (abs(fd) + 1e-3), (fd, Hb @ dH)


def _flow_e0_gen(o, T, gen):
    H = o.normal_order(T)
    HS, _, log, st = o.magnus_decouple(H, 0.5, REF, REF, 0, REF, generator=gen)
    return HS[0], st


@pytest.mark.parametrize("gen", [1, 2], ids=["atan", "w2"])
def test_flow_gradient_generators_vs_central_differences(gen):
    """the flow gradient with the atan and w2 generators (reverse sweep
    through their transposes) against central differences of the flow E0,
    eMax 2"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    T = synth.hamiltonian_vacuum(b, 0)
    e0, Tb, st = o.flow_grad(T, 0.5, REF, REF, REF, ckpt=4, generator=gen)
    e0_ref, st_ref = _flow_e0_gen(o, T, gen)
    assert st == st_ref and abs(e0 - e0_ref) < 1e-10
    h = 1e-5
    for s in range(3):
        d = _two_body_direction(b, 110 + s)
        ep, sp = _flow_e0_gen(o, T + h * d, gen)
        em, sm = _flow_e0_gen(o, T - h * d, gen)
        assert sp == sm == st
        fd = (ep - em) / (2 * h)
        assert abs(fd - Tb @ d) < 1e-6 * abs(fd), (fd, Tb @ d)
