"""Host-side basis of the product (imsrg_basis_create, no GPU needed) against
the reference's data files (committed fixture), the oracle's independent
layout construction, and the sizes derived from the reference's channel
rules (basic_IMSRG.f90:572-681, cross_coupled.f90:110-175)."""
import json
import os

import numpy as np
import pytest

from conftest import basis, oracle_basis

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# eMax: N orbits, channels, non-empty, max channel dim, two-body elements,
# CC blocks, max CC rows, max CC columns, WCC elements, max npp
SIZES = {
    4: (30, 60, 57, 84, 56_864, 40, 168, 26, 358_192, 61),
    6: (56, 84, 81, 265, 850_034, 56, 530, 42, 4_950_976, 232),
    8: (90, 108, 105, 648, 6_712_408, 72, 1_296, 58, 37_532_496, 606),
    10: (132, 132, 129, 1_341, 35_679_082, 88, 2_682, 74, 195_496_608, 1_291),
}


@pytest.mark.parametrize("emax", [4, 6, 8, 10])
def test_orbits_match_reference_sps(emax):
    from eom_imsrg import Basis
    g = json.load(open(os.path.join(GOLDEN, "hk_orbits.json")))
    b = basis(emax) if emax <= 8 else Basis.oxygen16(emax)
    tab = np.array(g[f"hk{emax}"])
    assert len(tab) == b.norb
    assert (tab[:, 0] == b.n).all() and (tab[:, 1] == b.l).all()
    assert (tab[:, 2] == b.jj).all() and (tab[:, 3] == b.tz).all()
    holes = np.array(g["holes_8_8"])
    assert (b.hole[: len(holes)] == holes[: b.norb]).all()
    assert b.hole[len(holes):].sum() == 0
    assert b.nhole == 6 and sum(b.jj[h] + 1 for h in b.holes) == 16


@pytest.mark.parametrize("emax", [4, 6, 8, 10])
def test_sizes(emax):
    from eom_imsrg import Basis, workload
    b = basis(emax) if emax <= 8 else Basis.oxygen16(emax)
    N, nch, ne, mx, e2, ncc, mr, mc, wcc, mpp = SIZES[emax]
    assert b.norb == N and b.nblocks == nch and int((b.dims.sum(1) > 0).sum()) == ne
    assert int(b.dims.sum(1).max()) == mx and b.e2_count() == e2
    assert b.ncc == ncc and int(b.rlen.max()) == mr and int(b.ncol.max()) == mc and b.wcclen == wcc
    assert int(b.dims[:, 0].max()) == mpp and int(b.dims[:, 2].max()) <= 5
    assert b.oplen == 1 + workload.e1(b) + e2


def test_workload_figures():
    """F_alg (18 pp/hh GEMM shapes + CC products + one-body GEMMs) and
    B_alg = 24 (E1 + E2) per commutator"""
    from eom_imsrg import Basis, workload
    ref_gf = {4: 0.0189, 6: 0.584, 8: 8.780, 10: 84.26}
    ref_mb = {4: 1.38, 6: 20.5, 8: 161.3, 10: 856.7}
    for e in ref_gf:
        b = basis(e) if e <= 8 else Basis.oxygen16(e)
        # the reference figures are quoted to 3-4 significant digits
        assert abs(workload.commutator_flops(b) / 1e9 - ref_gf[e]) < 0.0006 * ref_gf[e] + 5e-5
        assert abs(workload.commutator_bytes(b) / 1e6 - ref_mb[e]) < 0.003 * ref_mb[e]


@pytest.mark.parametrize("emax", [2, 4, 6])
def test_layout_matches_oracle(emax):
    b, o = basis(emax), oracle_basis(emax)
    assert (o.oplen, o.cclen, o.wcclen, o.ncc, o.nblocks) == (b.oplen, b.cclen, b.wcclen, b.ncc, b.nblocks)
    assert (o.dims == b.dims).all() and (o.goff == b.goff).all() and (o.jtp == b.jtp).all()
    for q in range(b.nblocks):
        for c in range(3):
            assert (o.pairs(q, c) == b.pairs[q][c]).all()
    rl, nc, co, wo, jtp = o.cc_layout()
    assert (rl == b.rlen).all() and (nc == b.ncol).all() and (co == b.ccoff).all() and (wo == b.wcoff).all()
    assert (jtp == b.ccjtp).all()


def test_channel_rules():
    """q = 3*(2J) + 2*(Tz+1) + P; pairs i <= j in orbit order, class pp/ph/hh,
    triangle, parity and Tz of the channel.  (a,a) pairs are listed at every
    J of the triangle, as the reference does; their odd-J elements are zero."""
    b = basis(4)
    for q in range(b.nblocks):
        J, T, P = (int(x) for x in b.jtp[q])
        assert q == 3 * J + 2 * (T + 1) + P        # J here is 2J
        for c in range(3):
            pr = b.pairs[q][c]
            assert len(pr) == b.dims[q, c]
            prev = (-1, -1)
            for i, j in pr:
                assert i <= j and (i, j) > prev
                prev = (i, j)
                assert b.hole[i] + b.hole[j] == c
                assert abs(b.jj[i] - b.jj[j]) <= J <= b.jj[i] + b.jj[j]
                assert (b.l[i] + b.l[j]) % 2 == P and b.tz[i] + b.tz[j] == 2 * T


def test_sixj_table_matches_oracle():
    """the product's cached 6j table (half-integer j1,j2,j4,j5; integer J3,J6)"""
    from oracle.oracle import sixj
    b = basis(4)
    n = 0
    for j1 in range(1, 10, 2):
        for j2 in range(1, 10, 2):
            for j4 in range(1, 10, 2):
                for j5 in range(1, 10, 2):
                    for J3 in range(0, 19, 2):
                        for J6 in range(0, 19, 2):
                            v = b.sixj(j1, j2, J3, j4, j5, J6)
                            assert abs(v - sixj(j1, j2, J3, j4, j5, J6)) < 1e-14
                            n += v != 0
    assert n > 1000
