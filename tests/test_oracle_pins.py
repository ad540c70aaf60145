"""Pins of the CPU oracle (oracle/imsrg_oracle.cpp) by checks that do not
reuse its production-order formulas:

* brute-force J-scheme commutators restated from brute_force_testing.f90:
  1765-2076 (orbit sums, 9j recoupling), group by group;
* angular-momentum closed forms, symmetries and orthogonality of 6j / 9j;
* [A,B] = -[B,A], the WCC = LCC RCC^T product against a plain loop;
* BCH: Omega = 0 identity, a one-body Omega as an exact orbital rotation of
  the vacuum operator followed by normal ordering, and the 1/n! structure
  from the eps^3 scaling of the truncated remainder;
* Magnus: Omega = 0 and Omega parallel to eta give dOmega = eta; the
  remainder after the 1, -1/2, 1/12 terms scales as eps^4;
* normal_order of the two-body identity (E0 = A(A-1)/2, f = A - n_p);
* build_gs_white on a Hamiltonian whose monopoles are known in closed form;
* commutator_220 of the white generator = 2 * mbpt2 when the Epstein-Nesbet
  and Moller-Plesset denominators coincide; mbpt2 against an orbit sum;
* the committed eMax-4 flow fixture.
All on the CPU, small bases.
"""
import json
import os

import numpy as np
import pytest
from scipy.linalg import expm

from conftest import assert_rel, basis, oracle_basis
from jscheme import (each_element, ljt_blocks, one_body, set_one_body, set_two_body, two_body_part)

HERMS = [(-1, 1), (-1, -1), (1, 1), (1, -1)]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rand_ops(b, hL, hR, seed):
    from eom_imsrg import synth
    mk = lambda h, s: synth.hermitian_random(b, s, 1.0) if h == 1 else synth.antihermitian(b, s, 1.0)
    return mk(hL, seed), mk(hR, seed + 1)


def _tri(a, b, c):
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


# ------------------------------------------------------------ 6j and 9j
def test_sixj_closed_form_with_zero():
    """{a b c; 0 c b} = (-1)^(a+b+c) / sqrt((2b+1)(2c+1)), arguments in 2j units"""
    from oracle.oracle import sixj
    n = 0
    for a in range(0, 9):
        for b in range(0, 9):
            for c in range(0, 9):
                if not _tri(a, b, c):
                    continue
                ref = (-1) ** ((a + b + c) // 2) / np.sqrt((b + 1) * (c + 1))
                assert abs(sixj(a, b, c, 0, c, b) - ref) < 1e-13, (a, b, c)
                n += 1
    assert n > 100


def test_sixj_symmetry_and_orthogonality():
    from oracle.oracle import sixj
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(3000):
        j = [int(x) for x in rng.integers(0, 10, 6)]
        j1, j2, j3, j4, j5, j6 = j
        if not (_tri(j1, j2, j3) and _tri(j1, j5, j6) and _tri(j4, j2, j6) and _tri(j4, j5, j3)):
            continue
        v = sixj(*j)
        # column permutations and exchange of upper / lower in two columns
        for p in [(j2, j1, j3, j5, j4, j6), (j1, j3, j2, j4, j6, j5), (j3, j2, j1, j6, j5, j4),
                  (j4, j5, j3, j1, j2, j6), (j1, j5, j6, j4, j2, j3)]:
            assert abs(sixj(*p) - v) < 1e-13
        checked += 1
    assert checked > 50
    # sum_x (2x+1)(2p+1) {a b x; c d p}{a b x; c d p'} = delta(p, p')
    for a, b, c, d in [(1, 3, 3, 1), (3, 5, 1, 3), (2, 4, 4, 2), (5, 5, 3, 3), (7, 3, 5, 5)]:
        ps = [p for p in range(0, 16) if _tri(a, d, p) and _tri(c, b, p)]
        xs = [x for x in range(0, 16) if _tri(a, b, x) and _tri(c, d, x)]
        for p in ps:
            for pp in ps:
                s = sum((x + 1) * (p + 1) * sixj(a, b, x, c, d, p) * sixj(a, b, x, c, d, pp) for x in xs)
                assert abs(s - (1.0 if p == pp else 0.0)) < 1e-12, (a, b, c, d, p, pp)


def test_ninej_reduces_to_sixj_and_symmetry():
    """{a b e; c d e; f f 0} = (-1)^(b+c+e+f) / sqrt((2e+1)(2f+1)) {a b e; d c f}"""
    from oracle.oracle import ninej, sixj
    n = 0
    for a in range(0, 7):
        for b in range(0, 7):
            for c in range(0, 7):
                for d in range(0, 7):
                    for e in range(0, 9):
                        for f in range(0, 9):
                            if not (_tri(a, b, e) and _tri(c, d, e) and _tri(a, c, f) and _tri(b, d, f)):
                                continue
                            ref = (-1) ** ((b + c + e + f) // 2) / np.sqrt((e + 1) * (f + 1)) * sixj(a, b, e, d, c, f)
                            assert abs(ninej(a, b, e, c, d, e, f, f, 0) - ref) < 1e-12
                            n += 1
    assert n > 200
    rng = np.random.default_rng(5)
    seen = 0
    for _ in range(60000):
        m = rng.integers(0, 8, 9)
        r = m.reshape(3, 3)
        ok = all(_tri(*r[i]) for i in range(3)) and all(_tri(*r[:, i]) for i in range(3))
        if not ok:
            continue
        v = ninej(*[int(x) for x in m])
        assert abs(ninej(*[int(x) for x in r.T.reshape(-1)]) - v) < 1e-12          # transpose
        sg = (-1) ** (int(m.sum()) // 2)                                             # odd row permutation
        assert abs(ninej(*[int(x) for x in r[[1, 0, 2]].reshape(-1)]) - sg * v) < 1e-12
        seen += 1
    assert seen > 20


# ------------------------------------------------------------ brute force
@pytest.mark.parametrize("hL,hR", HERMS)
def test_commutator_groups_vs_brute_force(hL, hR):
    """each group of the production-order commutator against the orbit-sum
    brute force of brute_force_testing.f90:1765-2076, every element at eMax 2"""
    b, o = basis(2), oracle_basis(2)
    L, R = _rand_ops(b, hL, hR, 41)
    hres = -hL * hR
    one = slice(1, b.off_fpp + b.npart * b.npart)
    two = two_body_part(b)

    # zero-body: 110 with one-body parts only, 220 with two-body parts only
    L1, R1, L2, R2 = L.copy(), R.copy(), L.copy(), R.copy()
    L1[two] = 0; R1[two] = 0; L2[one] = 0; R2[one] = 0
    assert abs(o.comm110(L1, hL, R1, hR) - o.bf_0body(L1, hL, R1, hR)) < 1e-11 * (1 + abs(o.bf_0body(L1, hL, R1, hR)))
    assert abs(o.comm220(L2, hL, R2, hR) - o.bf_0body(L2, hL, R2, hR)) < 1e-11 * (1 + abs(o.bf_0body(L2, hL, R2, hR)))
    assert abs(o.commutator(L, hL, R, hR, hres)[0] - o.bf_0body(L, hL, R, hR)) < 1e-10 * (1 + abs(o.bf_0body(L, hL, R, hR)))

    # one-body: 111 (one-body inputs), 121 (one x two), 221 (two-body inputs), all
    N = b.norb
    La = L.copy(); La[one] = 0          # two-body L
    Rb = R.copy(); Rb[two] = 0          # one-body R
    for (X, Y, what) in [(L1, R1, "111"), (La, Rb, "121"), (L2, R2, "221"), (L, R, "full")]:
        got = one_body(o, b, o.commutator(X, hL, Y, hR, hres), hres)
        ref = np.array([[o.bf_1body(X, hL, Y, hR, a, c) for c in range(N)] for a in range(N)])
        assert_rel(got, ref, what=f"one-body {what}")

    # two-body groups: 122 (1b x 2b), 222_pp_hh (ladder), Pandya + 222_ph (9j)
    r122 = np.zeros(b.oplen)
    o.comm122(L, hL, R, hR, r122, hres)
    rpp = np.zeros(b.oplen)
    o.comm222pphh(L, hL, R, hR, rpp, hres, np.zeros(b.oplen), np.zeros(b.oplen))
    rph = np.zeros(b.oplen)
    o.comm222ph(o.cross_coupled(L, hL), hL, o.cross_coupled(R, hR), hR, rph, hres)
    full = o.commutator(L, hL, R, hR, hres)
    got = {k: [] for k in ("122", "pphh", "ph", "all")}
    ref = {k: [] for k in got}
    size = []
    for q, g, ix, jx, (a, bb), (c, d), J in each_element(b):
        v, parts = o.bf_2body(L, hL, R, hR, a, bb, c, d, J)
        for k, arr, p in (("122", r122, parts[0]), ("pphh", rpp, parts[1]), ("ph", rph, parts[2]), ("all", full, v)):
            got[k].append(o.v_elem(arr, hres, a, bb, c, d, J))
            ref[k].append(p)
        size.append(float(np.abs(parts).sum()))
    assert len(ref["all"]) > 1000
    for k in ("122", "pphh", "ph"):
        assert_rel(got[k], ref[k], what=f"two-body {k}")
    # The 222_ph scatter uses the reference's literal 1/sqrt(2) = .70710678118
    # (commutators.f90:834,847; relative error -9.3e-12) where the brute force
    # has the exact sqrt(2): for i = j and k = l the ph group is 1.85e-11 low.
    # The sum of the groups is held to 1e-10 of the groups' magnitudes.
    d = np.abs(np.array(got["all"]) - np.array(ref["all"]))
    bound = 1e-10 * (np.array(size) + 1e-3 * max(size))
    assert (d <= bound).all(), f"two-body total: worst {float((d / bound).max()):.2f} of the bound"


def test_isolated_terms_sum_to_commutator():
    """the terms called one by one in the BCH order add up to o.commutator"""
    b, o = basis(2), oracle_basis(2)
    L, R = _rand_ops(b, -1, 1, 7)
    hres = 1
    RES = np.zeros(b.oplen)
    RES[0] = o.comm110(L, -1, R, 1) + o.comm220(L, -1, R, 1)
    o.comm111(L, -1, R, 1, RES, hres)
    o.comm121(L, -1, R, 1, RES, hres)
    RES2 = RES.copy()
    o.comm122(L, -1, R, 1, RES2, hres)
    RES[two_body_part(b)] = RES2[two_body_part(b)]
    w1, w2 = np.zeros(b.oplen), np.zeros(b.oplen)
    o.comm222pphh(L, -1, R, 1, RES, hres, w1, w2)
    o.comm221(L, -1, R, 1, RES, hres, w1, w2)
    o.comm222ph(o.cross_coupled(L, -1), -1, o.cross_coupled(R, 1), 1, RES, hres)
    assert_rel(RES, o.commutator(L, -1, R, 1, hres), rtol=1e-13, what="term sum")


@pytest.mark.parametrize("emax", [2, 4])
def test_commutator_antisymmetric(emax):
    b, o = basis(emax), oracle_basis(emax)
    for hL, hR in HERMS:
        L, R = _rand_ops(b, hL, hR, 17)
        hres = -hL * hR
        AB = o.commutator(L, hL, R, hR, hres)
        BA = o.commutator(R, hR, L, hL, hres)
        assert_rel(AB, -BA, rtol=1e-12, what=f"[A,B]+[B,A] herm {hL},{hR}")


def test_ph_product_against_loop():
    """WCC(q) = LCC(q) RCC(q)^T, checked with an unoptimised einsum loop"""
    b, o = basis(2), oracle_basis(2)
    L, R = _rand_ops(b, -1, 1, 23)
    LCC, RCC = o.cross_coupled(L, -1), o.cross_coupled(R, 1)
    W = o.comm222ph(LCC, -1, RCC, 1, np.zeros(b.oplen), 1)
    for q in range(b.ncc):
        r, nb = int(b.rlen[q]), int(b.ncol[q])
        if r == 0:
            continue
        Lq = LCC[b.ccoff[q]: b.ccoff[q] + r * nb].reshape(r, nb, order="F")
        Rq = RCC[b.ccoff[q]: b.ccoff[q] + r * nb].reshape(r, nb, order="F")
        Wq = W[b.wcoff[q]: b.wcoff[q] + r * r].reshape(r, r, order="F")
        ref = np.einsum("ik,jk->ij", Lq, Rq, optimize=False)
        assert_rel(Wq, ref, rtol=1e-13, what=f"WCC block {q}", scale=max(float(np.abs(ref).max()), 1e-300))


# ------------------------------------------------------------ BCH
def test_bch_zero_generator_is_identity():
    b, o = basis(4), oracle_basis(4)
    from eom_imsrg import synth
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    HS, it, _ = o.bch(np.zeros(b.oplen), H, 1e-10)
    assert np.array_equal(HS, H)
    assert it >= 1


def _vacuum_rotation(b, o, H0, U):
    """e^Omega H0 e^-Omega for a one-body Omega = log U mixing equal-(l,j,tz)
    orbits: t' = U t U^T, V'_abcd = sum U_aa' U_bb' U_cc' U_dd' V_a'b'c'd'"""
    N = b.norb
    t = one_body(o, b, H0, 1)
    out = np.zeros(b.oplen)
    set_one_body(b, out, U @ t @ U.T)
    nz = [np.nonzero(np.abs(U[a]) > 0)[0] for a in range(N)]
    cache = {}

    def v0(a, bb, c, d, J):
        k = (a, bb, c, d, J)
        if k not in cache:
            cache[k] = o.v_elem(H0, 1, a, bb, c, d, J)
        return cache[k]

    def rotated(a, bb, c, d, J):
        s = 0.0
        for a1 in nz[a]:
            for b1 in nz[bb]:
                for c1 in nz[c]:
                    for d1 in nz[d]:
                        s += U[a, a1] * U[bb, b1] * U[c, c1] * U[d, d1] * v0(a1, b1, c1, d1, J)
        return s

    set_two_body(b, out, rotated)
    return out


def test_bch_one_body_generator_is_orbital_rotation():
    """BCH(Omega1, NO(H0)) = NO(U H0 U^-1) exactly: a one-body generator
    keeps the commutators inside the (0,1,2)-body space"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H0 = synth.hamiltonian_vacuum(b, 0)
    Om = synth.antihermitian(b, 9, 0.3)
    Om[two_body_part(b)] = 0
    W = one_body(o, b, Om, -1)
    assert np.allclose(W, -W.T) and np.abs(W).max() > 0.1
    U = np.eye(b.norb)
    for blk in ljt_blocks(b):
        U[np.ix_(blk, blk)] = expm(W[np.ix_(blk, blk)])
    ref = o.normal_order(_vacuum_rotation(b, o, H0, U))
    HS, it, _ = o.bch(Om, o.normal_order(H0), 1e-13)
    assert it > 5
    assert_rel(HS, ref, rtol=1e-10, what="BCH orbital rotation")


def test_bch_truncation_scales_as_eps_cubed():
    """HS(eps) - H - eps[O,H] - eps^2/2 [O,[O,H]] = O(eps^3)"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    Om = synth.antihermitian(b, 3, 0.05)
    C1 = o.commutator(Om, -1, H, 1, 1)
    C2 = o.commutator(Om, -1, C1, 1, 1)
    res = []
    for eps in (0.4, 0.2, 0.1):
        HS, _, _ = o.bch(eps * Om, H, 1e-14)
        res.append(np.linalg.norm(HS - H - eps * C1 - 0.5 * eps * eps * C2))
    for r0, r1 in zip(res, res[1:]):
        assert 7.0 < r0 / r1 < 9.0, res


# ------------------------------------------------------------ Magnus
def test_magnus_zero_and_parallel_generator():
    from eom_imsrg import synth
    b, o = basis(4), oracle_basis(4)
    eta = synth.antihermitian(b, 4, 0.05)
    out, _ = o.magnus(eta, np.zeros(b.oplen), 1e-10)
    assert np.array_equal(out, eta)
    out, _ = o.magnus(eta, 1e-12 * eta, 1e-10)     # ||Omega|| < 1e-9: dOmega = eta
    assert np.array_equal(out, eta)
    out, _ = o.magnus(0.7 * eta, 2.0 * eta, 1e-10)  # [Omega, eta] = 0
    assert_rel(out, 0.7 * eta, rtol=1e-10, what="Magnus, Omega parallel to eta", scale=0.7 * np.abs(eta).max())


def test_magnus_truncation_scales_as_eps_fourth():
    """dOmega(eps) - eta + eps/2 [O,eta] - eps^2/12 [O,[O,eta]] = O(eps^4)
    (the eps^3 coefficient of the series is zero)"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    eta = synth.antihermitian(b, 4, 0.05)
    Om = synth.antihermitian(b, 6, 0.05)
    C1 = o.commutator(Om, -1, eta, -1, -1)
    C2 = o.commutator(Om, -1, C1, -1, -1)
    res = []
    for eps in (1.6, 0.8, 0.4):
        out, n = o.magnus(eta, eps * Om, 1e-15)
        res.append(np.linalg.norm(out - eta + 0.5 * eps * C1 - eps * eps / 12.0 * C2))
    for r0, r1 in zip(res, res[1:]):
        assert 13.0 < r0 / r1 < 19.0, res


# ------------------------------------------------------------ normal order, generator, mbpt2
def _identity_two_body(b, op):
    """V = 1 on every stored pair (Pauli-forbidden (a,a) at odd J excluded)"""
    for q, g, ix, jx, (a, bb), (c, d), J in each_element(b):
        if g in (0, 3, 4) and ix == jx and not (a == bb and (J // 2) % 2 == 1):
            n1, _ = b.block_shape(q, g)
            op[int(b.goff[q, g]) + ix + jx * n1] = 1.0


@pytest.mark.parametrize("emax", [2, 4])
def test_normal_order_two_body_identity(emax):
    """V = 1 is (1/2) N(N-1): E0 = A(A-1)/2, f_pp = A - n_p, V unchanged"""
    b, o = basis(emax), oracle_basis(emax)
    H = np.zeros(b.oplen)
    _identity_two_body(b, H)
    out = o.normal_order(H)
    A = int(sum(b.jj[h] + 1 for h in b.holes))
    assert A == 16
    assert abs(out[0] - A * (A - 1) / 2) < 1e-11
    F = one_body(o, b, out, 1)
    ref = np.diag([A - 1.0 if b.hole[p] else float(A) for p in range(b.norb)])
    assert np.abs(F - ref).max() < 1e-12
    assert np.array_equal(out[two_body_part(b)], H[two_body_part(b)])
    # one-body vacuum operator: E0 = sum_h (2j_h+1) t_hh, f = t
    from eom_imsrg import synth
    T = synth.hamiltonian_vacuum(b, 1)
    T[two_body_part(b)] = 0
    out = o.normal_order(T)
    assert abs(out[0] - sum((b.jj[h] + 1) * o.f_elem(T, 1, h, h) for h in b.holes)) < 1e-11
    assert np.array_equal(out[1:], T[1:])


def test_white_generator_closed_form():
    """with V = 1 + X (X in pphh only) the Epstein-Nesbet denominators are
    f_a+f_b-f_i-f_j + m(ab) + m(ij) - 4, with m(xy) = 1 for x != y and
    2j/(2j+1) for x = y; the ph one is f_aa - f_ii - 1"""
    b, o = basis(4), oracle_basis(4)
    rng = np.random.default_rng(8)
    H = np.zeros(b.oplen)
    N = b.norb
    F = np.diag([(-30.0 if b.hole[a] else 20.0) + rng.normal() for a in range(N)])
    for blk in ljt_blocks(b):
        for a in blk:
            for i in blk:
                if not b.hole[a] and b.hole[i]:
                    F[a, i] = F[i, a] = rng.normal()
    set_one_body(b, H, F)
    _identity_two_body(b, H)
    for q in range(b.nblocks):
        n1, n2 = b.block_shape(q, 2)
        o2 = int(b.goff[q, 2])
        H[o2: o2 + n1 * n2] = rng.normal(size=n1 * n2)
    eta = o.white(H)
    E = one_body(o, b, eta, -1)
    for a in b.parts:
        for i in b.holes:
            ref = F[a, i] / (F[a, a] - F[i, i] - 1.0) if F[a, i] != 0 else 0.0
            assert abs(E[a, i] - ref) < 1e-13 * (1 + abs(ref))
    assert np.abs(E[np.ix_(b.parts, b.parts)]).max() == 0 and np.abs(E[np.ix_(b.holes, b.holes)]).max() == 0
    m = lambda x, y: 1.0 if x != y else b.jj[x] / (b.jj[x] + 1.0)
    n = 0
    for q, g, ix, jx, (a, bb), (i, j), J in each_element(b):
        n1, _ = b.block_shape(q, g)
        k = int(b.goff[q, g]) + ix + jx * n1
        if g != 2:
            assert eta[k] == 0.0
            continue
        den = F[a, a] + F[bb, bb] - F[i, i] - F[j, j] + m(a, bb) + m(i, j) - 4.0
        assert abs(eta[k] - H[k] / den) < 1e-13 * (1 + abs(H[k] / den))
        n += 1
    assert n > 100


def _closed_form_hamiltonian(b, seed):
    """H of the closed-form white test: returns (H, F, denominator(q, g, ix, jx) list)"""
    rng = np.random.default_rng(seed)
    H = np.zeros(b.oplen)
    N = b.norb
    F = np.diag([(-30.0 if b.hole[a] else 20.0) + rng.normal() for a in range(N)])
    for blk in ljt_blocks(b):
        for a in blk:
            for i in blk:
                if not b.hole[a] and b.hole[i]:
                    F[a, i] = F[i, a] = 3.0 * rng.normal()
    set_one_body(b, H, F)
    _identity_two_body(b, H)
    for q in range(b.nblocks):
        n1, n2 = b.block_shape(q, 2)
        o2 = int(b.goff[q, 2])
        H[o2: o2 + n1 * n2] = 5.0 * rng.normal(size=n1 * n2)
    return H, F


def _check_atan_form(b, o, eta, Hn, F, what):
    """eta = 0.5 atan(2 Hn / Eden) with the closed-form denominators of the
    V = 1 + X Hamiltonian (F its one-body matrix), zero outside fph / gam3"""
    m = lambda x, y: 1.0 if x != y else b.jj[x] / (b.jj[x] + 1.0)
    E = one_body(o, b, eta, -1)
    Fn = one_body(o, b, Hn, 1)
    for a in b.parts:
        for i in b.holes:
            ref = 0.5 * np.arctan(2 * Fn[a, i] / (F[a, a] - F[i, i] - 1.0)) if Fn[a, i] != 0 else 0.0
            assert abs(E[a, i] - ref) < 1e-13 * (1 + abs(ref)), what
    assert np.abs(E[np.ix_(b.parts, b.parts)]).max() == 0 and np.abs(E[np.ix_(b.holes, b.holes)]).max() == 0
    n = 0
    for q, g, ix, jx, (a, bb), (i, j), J in each_element(b):
        n1, _ = b.block_shape(q, g)
        k = int(b.goff[q, g]) + ix + jx * n1
        if g != 2:
            assert eta[k] == 0.0, what
            continue
        den = F[a, a] + F[bb, bb] - F[i, i] - F[j, j] + m(a, bb) + m(i, j) - 4.0
        ref = 0.5 * np.arctan(2 * Hn[k] / den)
        assert abs(eta[k] - ref) < 1e-13 * (1 + abs(ref)), what
        n += 1
    assert n > 100
    return n


def test_atan_generator_closed_form():
    """build_gs_atan: 0.5 atan(2 H / Eden) on fph and the pphh blocks with
    the closed-form Epstein-Nesbet denominators; couplings up to several
    times the gap so atan is far from linear"""
    b, o = basis(4), oracle_basis(4)
    H, F = _closed_form_hamiltonian(b, 31)
    eta = o.gs_atan(H)
    _check_atan_form(b, o, eta, H, F, "atan")
    big = np.abs(2 * H[1:] / 10.0).max()
    assert big > 1.0


def test_atanxx_takes_denominators_from_hold():
    """build_gs_atanxx(Hold, H): denominators of Hold (closed form), numerators
    of an unrelated H; atanxx(H, H) = atan(H) to the bit"""
    b, o = basis(4), oracle_basis(4)
    Hold, F = _closed_form_hamiltonian(b, 32)
    from eom_imsrg import synth
    Hn = o.normal_order(synth.hamiltonian_vacuum(b, 3))
    eta = o.gs_atanxx(Hold, Hn)
    _check_atan_form(b, o, eta, Hn, F, "atanxx")
    assert np.array_equal(o.gs_atanxx(Hn, Hn), o.gs_atan(Hn))


def test_w2_generator_composition_and_small_coupling():
    """build_gs_w2 = 1.5 atan(H) + 0.5 atanxx(H, [atan(H), H]) from the
    separately pinned generators and commutator, E0 = 0.  With Moller-Plesset
    denominators (diagonal f, V only in pphh) and V scaled by t, [eta, f]
    = -V exactly and [eta, V] has no pphh part, so w2 = 1.5 white - 0.5
    white + O(t^3) (atan's cubic term): the relative deviation falls as t^2"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 5))
    e1 = o.gs_atan(H)
    H2 = o.commutator(e1, -1, H, 1, 1)
    ref = 1.5 * e1 + 0.5 * o.gs_atanxx(H, H2)
    ref[0] = 0.0
    w = o.gs_w2(H)
    assert w[0] == 0.0
    assert_rel(w, ref, rtol=1e-13, what="w2 composition")
    Hmp = _mp_hamiltonian(b, 14)
    r = []
    for t in (2e-3, 1e-3):
        Ht = Hmp.copy()
        for q in range(b.nblocks):
            n1, n2 = b.block_shape(q, 2)
            o2 = int(b.goff[q, 2])
            Ht[o2: o2 + n1 * n2] *= t
        wt = o.white(Ht)
        d = o.gs_w2(Ht)[1:] - wt[1:]
        r.append(np.abs(d).max() / np.abs(wt[1:]).max())
    assert r[0] < 1e-7 and 0.2 < r[1] / r[0] < 0.3, r


def _mp_hamiltonian(b, seed):
    """diagonal f, zero ph one-body, V only in the pphh blocks: Epstein-Nesbet
    = Moller-Plesset denominators, no 110 term"""
    rng = np.random.default_rng(seed)
    H = np.zeros(b.oplen)
    set_one_body(b, H, np.diag([(-25.0 if b.hole[a] else 15.0) + 3 * rng.normal() for a in range(b.norb)]))
    for q in range(b.nblocks):
        n1, n2 = b.block_shape(q, 2)
        o2 = int(b.goff[q, 2])
        H[o2: o2 + n1 * n2] = rng.normal(size=n1 * n2)
    return H


@pytest.mark.parametrize("emax", [2, 4])
def test_white_220_is_twice_mbpt2(emax):
    b, o = basis(emax), oracle_basis(emax)
    H = _mp_hamiltonian(b, 12)
    eta = o.white(H)
    e2 = o.mbpt2(H)
    assert e2 < 0
    assert abs(o.comm110(eta, -1, H, 1)) < 1e-14
    assert abs(o.comm220(eta, -1, H, 1) - 2.0 * e2) < 1e-12 * abs(e2)
    assert abs(o.commutator(eta, -1, H, 1, 1)[0] - 2.0 * e2) < 1e-12 * abs(e2)


def test_mbpt2_orbit_sum():
    """E2 = 1/4 sum_{ab in p, ij in h, J} (2J+1) |V^J_abij|^2 / (f_i+f_j-f_a-f_b)"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    f = [o.f_elem(H, 1, a, a) for a in range(b.norb)]
    s = 0.0
    for a in b.parts:
        for bb in b.parts:
            for i in b.holes:
                for j in b.holes:
                    if b.tz[a] + b.tz[bb] != b.tz[i] + b.tz[j] or (b.l[a] + b.l[bb] + b.l[i] + b.l[j]) % 2:
                        continue
                    lo = max(abs(b.jj[a] - b.jj[bb]), abs(b.jj[i] - b.jj[j]))
                    hi = min(b.jj[a] + b.jj[bb], b.jj[i] + b.jj[j])
                    for J in range(lo, hi + 1, 2):
                        v = o.v_elem(H, 1, int(a), int(bb), int(i), int(j), J)
                        s += (J + 1) * v * v / (f[i] + f[j] - f[a] - f[bb])
    s *= 0.25
    assert abs(o.mbpt2(H) - s) < 1e-12 * abs(s)


# ------------------------------------------------------------ committed flow fixture
def test_oracle_flow_matches_fixture():
    from eom_imsrg import synth
    g = json.load(open(os.path.join(GOLDEN, "oracle_emax4.json")))
    b, o = basis(4), oracle_basis(4)
    tol = float(np.float32(1e-6))
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    assert abs(H[0] - g["E0_bare_normal_ordered"]) < 1e-9
    assert abs(o.mbpt2(H) - g["mbpt2_bare"]) < 1e-9
    G = synth.antihermitian(b, seed=1, sigma=0.05)
    C = o.commutator(G, -1, H, 1, 1)
    assert abs(C[0] - g["commutator_Omega_H"]["E0"]) < 1e-10
    assert abs(o.frob_norm(C) - g["commutator_Omega_H"]["norm"]) < 1e-9
    HS, _, log, st = o.magnus_decouple(H, 0.5, tol, tol, 0, tol)
    assert st == g["flow"]["steps"]
    assert [int(x) for x in log[:, 3]] == g["flow"]["commutators"]
    assert np.abs(log[:, 1] - np.array(g["flow"]["E0"])).max() < 1e-8
    assert abs(HS[0] - g["flow"]["final_E0"]) < 1e-8


def test_generators_flow_to_the_same_ground_state():
    """white, atan and w2 flows of one Hamiltonian reach |mbpt2| < 1e-6 and
    agree on E0 to the size the IMSRG(2) truncation and that tolerance leave
    (1e-4 MeV here), while a different Hamiltonian differs by MeV"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    tol = float(np.float32(1e-6))
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    E = []
    for gen in (0, 1, 2):
        HS, _, log, st = o.magnus_decouple(H, 0.5, tol, tol, 0, tol, generator=gen)
        assert 0 < st < 200 and abs(log[-1, 2] - log[-1, 1]) < tol
        E.append(HS[0])
    assert max(E) - min(E) < 1e-4, E
    assert E[0] != E[1] and E[0] != E[2]
