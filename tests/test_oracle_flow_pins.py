"""Pins of the oracle's mat_frob_norm and magnus_decouple loop by checks the
mathematics fixes, independent of the oracle's own code for them (CPU):

* mat_frob_norm (basic_IMSRG.f90:4056-4087) is the Frobenius norm of the
  full hermitian (or antihermitian) one-body matrix plus that of every full
  two-body channel matrix: the stored rectangular blocks stand for two
  blocks of the full matrix and the square blocks for one.  Checked against
  dense matrices assembled with numpy from the storage blocks, against the
  all-ones closed form N^2 + sum_q n_q^2, and with a single nonzero element
  in each of the nine block classes (weights 1, 2, 1 and 1, 2, 2, 1, 1, 2).
* magnus_decouple (IMSRG_MAGNUS.f90:15-128): every step k of the oracle's
  loop is recomposed from the separately pinned primitives (white, Magnus,
  Euler, BCH from the BARE H, mbpt2): Omega_1 = ds * eta(H), Omega_{k+1} =
  Omega_k + ds * MAGNUS(eta(HS_k), Omega_k), HS_k = BCH(Omega_k, H), the
  logged s, E0, E0 + mbpt2 and commutator counts, and the stop at the first
  |mbpt2| <= tol.  Two planted mistakes (BCH from H(s) instead of the bare
  H; the Euler step with eta instead of dOmega) are shown to fail these
  checks.  For a Hamiltonian without a two-body part the IMSRG(2) flow is an
  exact unitary orbital rotation, so the flowed E0 must reach the closed
  form sum over the hole blocks of (2j+1) * lowest eigenvalue.
"""
import numpy as np
import pytest

from conftest import assert_rel, basis, oracle_basis
from jscheme import ljt_blocks

TOL = float(np.float32(1e-6))   # the reference's conv / crit literal (default real)
DS = 0.5


# ------------------------------------------------------------ mat_frob_norm
def _dense_one_body(b, op, herm):
    A, P = b.nhole, b.npart
    F = np.zeros((b.norb, b.norb))
    fhh = op[b.off_fhh: b.off_fhh + A * A].reshape(A, A, order="F")
    fph = op[b.off_fph: b.off_fph + P * A].reshape(P, A, order="F")
    fpp = op[b.off_fpp: b.off_fpp + P * P].reshape(P, P, order="F")
    H, Pp = b.holes, b.parts
    F[np.ix_(H, H)] = fhh
    F[np.ix_(Pp, Pp)] = fpp
    F[np.ix_(Pp, H)] = fph
    F[np.ix_(H, Pp)] = herm * fph.T
    return F


def _dense_channel(b, op, herm, q):
    """full channel matrix over the pairs (pp, ph, hh): the lower blocks are
    herm times the transposes of the stored upper ones"""
    np_, nb, nh = (int(x) for x in b.dims[q])
    n = np_ + nb + nh
    V = np.zeros((n, n))
    r0 = (0, 0, 0, np_, np_ + nb, np_)
    c0 = (0, np_, np_ + nb, np_, np_ + nb, np_ + nb)
    for g in range(6):
        n1, n2 = b.block_shape(q, g)
        if n1 * n2 == 0:
            continue
        X = op[int(b.goff[q, g]): int(b.goff[q, g]) + n1 * n2].reshape(n1, n2, order="F")
        V[r0[g]: r0[g] + n1, c0[g]: c0[g] + n2] = X
        if g in (1, 2, 5):
            V[c0[g]: c0[g] + n2, r0[g]: r0[g] + n1] = herm * X.T
    return V


@pytest.mark.parametrize("emax,herm", [(2, 1), (4, -1), (4, 1)])
def test_frob_norm_is_the_norm_of_the_full_matrices(emax, herm):
    from eom_imsrg import synth
    b, o = basis(emax), oracle_basis(emax)
    op = synth.hermitian_random(b, 5, 1.0) if herm == 1 else synth.antihermitian(b, 5, 1.0)
    op[0] = 123.0   # E0 does not enter the norm
    ref = np.linalg.norm(_dense_one_body(b, op, herm)) ** 2
    ref += sum(np.linalg.norm(_dense_channel(b, op, herm, q)) ** 2 for q in range(b.nblocks))
    assert abs(o.frob_norm(op) - np.sqrt(ref)) < 1e-12 * np.sqrt(ref)


@pytest.mark.parametrize("emax", [2, 4])
def test_frob_norm_all_ones_closed_form(emax):
    b, o = basis(emax), oracle_basis(emax)
    op = np.ones(b.oplen)
    n_q = b.dims.sum(axis=1).astype(np.float64)
    ref = np.sqrt(float(b.norb) ** 2 + float(np.sum(n_q * n_q)))
    assert abs(o.frob_norm(op) - ref) < 1e-12 * ref


def test_frob_norm_single_element_per_block_class():
    """one nonzero stored element v: norm = |v| sqrt(w), w = 2 exactly for
    the blocks that stand for two blocks of the full matrix"""
    b, o = basis(4), oracle_basis(4)
    A, P = b.nhole, b.npart
    rng = np.random.default_rng(11)
    v = -3.25
    starts = [(b.off_fhh, A * A, 1.0), (b.off_fph, P * A, 2.0), (b.off_fpp, P * P, 1.0)]
    weights = (1.0, 2.0, 2.0, 1.0, 1.0, 2.0)
    for g in range(6):
        q = next(q for q in range(b.nblocks) if np.prod(b.block_shape(q, g)) > 0)
        starts.append((int(b.goff[q, g]), int(np.prod(b.block_shape(q, g))), weights[g]))
    for off, n, w in starts:
        op = np.zeros(b.oplen)
        op[off + int(rng.integers(0, n))] = v
        assert abs(o.frob_norm(op) - abs(v) * np.sqrt(w)) < 1e-14, (off, n, w)


# ------------------------------------------------------------ magnus_decouple
def _truncated_states(o, H, nsteps):
    """(HS_k, Omega_k) after k = 1..nsteps steps of the oracle's loop"""
    out = []
    for k in range(1, nsteps + 1):
        HS, G, _, st = o.magnus_decouple(H, DS, TOL, TOL, 0, TOL, maxsteps=k)
        assert st in (k, -k)
        out.append((HS, G))
    return out


def _loop_defects(o, H, states):
    """largest violation of each loop identity over the steps"""
    d = {"omega1": 0.0, "bch_from_bare": 0.0, "euler": 0.0}
    eta0 = o.white(H)
    d["omega1"] = float(np.abs(states[0][1] - DS * eta0).max())
    for k, (HS, G) in enumerate(states):
        ref, _, _ = o.bch(G, H, TOL)
        d["bch_from_bare"] = max(d["bch_from_bare"], float(np.abs(HS - ref).max() / np.abs(ref).max()))
        if k + 1 < len(states):
            dG, _ = o.magnus(o.white(HS), G, TOL)
            nxt = G + DS * dG
            d["euler"] = max(d["euler"], float(np.abs(states[k + 1][1] - nxt).max() / np.abs(nxt).max()))
    return d


def _planted_states(o, H, nsteps, mistake):
    """the loop with one planted mistake, composed from the oracle's primitives"""
    HS, G = H.copy(), np.zeros_like(H)
    DG = o.white(HS)
    out = []
    for _ in range(nsteps):
        dG, _ = o.magnus(DG, G, TOL)
        G = G + DS * (DG if mistake == "euler_eta" else dG)
        HS, _, _ = o.bch(G, HS if mistake == "bch_from_hs" else H, TOL)
        DG = o.white(HS)
        out.append((HS, G))
    return out


def test_magnus_decouple_loop_identities():
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    HS, G, log, n = o.magnus_decouple(H, DS, TOL, TOL, 0, TOL)
    assert 5 <= n < 200
    states = _truncated_states(o, H, n)
    assert np.array_equal(states[-1][0], HS) and np.array_equal(states[-1][1], G)
    d = _loop_defects(o, H, states)
    assert d["omega1"] == 0.0, d                       # dOmega = eta exactly while ||Omega|| = 0
    assert d["bch_from_bare"] < 1e-13 and d["euler"] < 1e-13, d
    # the log and the stop rule
    for k, (HSk, Gk) in enumerate(states):
        m = o.mbpt2(HSk)
        assert abs(log[k, 0] - (k + 1) * DS) < 1e-15
        assert log[k, 1] == HSk[0] and abs(log[k, 2] - (HSk[0] + m)) < 1e-12
        assert (abs(m) > TOL) == (k + 1 < n), (k, m)
        prev = states[k - 1][0] if k else H
        prevG = states[k - 1][1] if k else np.zeros_like(H)
        _, nm = o.magnus(o.white(prev), prevG, TOL)
        _, nb, _ = o.bch(Gk, H, TOL)
        assert int(log[k, 3]) == nm + nb - 1
    assert abs(o.mbpt2(H)) > TOL


def test_magnus_decouple_planted_mistakes_fail_the_identities():
    """the same identities reject a loop that takes BCH from H(s) or steps
    Omega with eta: the checks above can tell a wrong loop"""
    from eom_imsrg import synth
    b, o = basis(2), oracle_basis(2)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    d = _loop_defects(o, H, _planted_states(o, H, 4, "bch_from_hs"))
    assert d["bch_from_bare"] > 1e-6, d
    d = _loop_defects(o, H, _planted_states(o, H, 4, "euler_eta"))
    assert d["euler"] > 1e-6, d


def test_magnus_decouple_one_body_flow_reaches_exact_energy():
    """H without a two-body part: every commutator stays one-body, BCH is an
    exact unitary rotation of the orbitals, and the white flow decouples the
    hole space, so E0 -> sum over the (l,j,tz) blocks of (2j+1) times the
    block's lowest n_hole eigenvalues of the one-body matrix"""
    from eom_imsrg import synth
    b, o = basis(4), oracle_basis(4)
    Hv = synth.hamiltonian_vacuum(b, seed=3, hw=20.0)
    Hv[b.off_fpp + b.npart * b.npart:] = 0.0
    H = o.normal_order(Hv)
    F = _dense_one_body(b, H, 1)
    exact = 0.0
    for blk in ljt_blocks(b):
        nh = int(sum(b.hole[a] for a in blk))
        if nh:
            ev = np.linalg.eigvalsh(F[np.ix_(blk, blk)])
            exact += (b.jj[blk[0]] + 1) * float(np.sum(ev[:nh]))
    assert abs(H[0] - exact) > 1e-3          # the reference state is not yet the ground state
    HS, G, log, n = o.magnus_decouple(H, DS, 1e-13, TOL, 1, 1e-12, maxsteps=200)
    assert 0 < n < 200
    assert abs(HS[0] - exact) < 1e-11 * abs(exact), (HS[0], exact)
    assert np.abs(HS[b.off_fph: b.off_fph + b.npart * b.nhole]).max() < 1e-5
    assert np.all(HS[b.off_fpp + b.npart * b.npart:] == 0.0)
