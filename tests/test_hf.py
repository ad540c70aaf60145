"""Hartree-Fock normal ordering (calc_HF, HF_mod.f90:14-105, without
three-body forces): the SCF loop, the one-body and two-body basis
transforms and the HF energy.

The oracle's restatement is pinned by
* the two-body transform against the four-index sum
  V'(ab,cd) = sum D(i,a) D(j,b) D(k,c) D(l,d) V(ij,kl) over v_elem, for
  random orthogonal rotations of every (l, j, tz) block;
* the HF Hamiltonian against normal_order of the rotated vacuum Hamiltonian
  (E0 = the HF energy, f = the Fock matrix in the HF basis, Gamma = the
  rotated two-body part), to the SCF tolerance;
* the Brillouin condition (the particle-hole Fock elements vanish) and the
  variational bound E_HF <= E0 of the HO reference.
The GPU test compares the library's HF (host SCF, two-body transform as two
grouped MFMA GEMM families) with the oracle.
"""
import numpy as np
import pytest

from conftest import basis, oracle_basis


def _rotations(ob, seed):
    sz, off = ob.sp_blocks()
    rng = np.random.default_rng(seed)
    D = np.zeros(int(off[-1]))
    for q, n in enumerate(sz):
        if n:
            Q, R = np.linalg.qr(rng.standard_normal((n, n)))
            D[off[q]: off[q + 1]] = (Q * np.sign(np.diag(R))).T.reshape(-1)
    return D


def _dsmall(ob, b, D):
    """Dsmall(a, i): component of old orbit a in new orbit i"""
    sz, off = ob.sp_blocks()
    N = b.norb
    blocks = []
    for t in (-1, 1):
        for l in range(int(b.l.max()) + 1):
            for j in range(1, int(b.jj.max()) + 1, 2):
                blocks.append([i for i in range(N) if b.l[i] == l and b.jj[i] == j and b.tz[i] == t])
    assert [len(x) for x in blocks] == list(sz)
    Ds = np.zeros((N, N))
    for q, st in enumerate(blocks):
        n = len(st)
        M = D[off[q]: off[q + 1]].reshape(n, n).T
        for x in range(n):
            for y in range(n):
                Ds[st[x], st[y]] = M[x, y]
    return Ds, blocks


def test_transform_2b_four_index():
    """every stored element at eMax 2 and a sample at eMax 4"""
    from eom_imsrg import synth
    for emax, nsample in ((2, None), (4, 400)):
        b, ob = basis(emax), oracle_basis(emax)
        H = synth.hamiltonian_vacuum(b, seed=2)
        D = _rotations(ob, 7)
        Hp = ob.hf_transform_2b(H, D)
        Ds, blocks = _dsmall(ob, b, D)
        blk = {o: st for st in blocks for o in st}
        items = []
        for q in range(ob.nblocks):
            J = int(ob.jtp[q, 0])
            for g, (cr, cc) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (2, 2), (1, 2))):
                for (a, bb) in ob.pairs(q, cr):
                    for (c, d) in ob.pairs(q, cc):
                        items.append((a, bb, c, d, J))
        rng = np.random.default_rng(1)
        if nsample:
            items = [items[x] for x in rng.choice(len(items), nsample, replace=False)]
        for (a, bb, c, d, J) in items:
            want = 0.0
            for i in blk[a]:
                for j in blk[bb]:
                    for k in blk[c]:
                        for l in blk[d]:
                            w = Ds[i, a] * Ds[j, bb] * Ds[k, c] * Ds[l, d]
                            if w:
                                want += w * ob.v_elem(H, 1, i, j, k, l, J)
            got = ob.v_elem(Hp, 1, a, bb, c, d, J)
            assert abs(got - want) < 1e-11 * (1 + abs(want)), (emax, a, bb, c, d, J, got, want)


@pytest.mark.parametrize("emax", [2, 4])
def test_hf_is_normal_ordered_rotated_vacuum(emax):
    """calc_HF's result = normal_order of (D^T t D, transformed V); the HF
    energy and Fock matrix lag the final rotation by one SCF update, so the
    agreement is to the SCF tolerance"""
    from eom_imsrg import synth
    b, ob = basis(emax), oracle_basis(emax)
    H = synth.hamiltonian_vacuum(b, seed=4, hw=20.0)
    Hhf, D, it = ob.hartree_fock(H)
    assert 3 <= it < 100
    # rotate the vacuum: one-body blocks by D^T t D, two-body by the transform
    Ds, blocks = _dsmall(ob, b, D)
    Hrot = ob.hf_transform_2b(H, D)
    t = np.array([[ob.f_elem(H, 1, x, y) for y in range(b.norb)] for x in range(b.norb)])
    tr = Ds.T @ t @ Ds
    Hrot[1: b.off_fpp + b.npart * b.npart] = 0.0
    for x in range(b.norb):
        for y in range(b.norb):
            if abs(tr[x, y]) > 0 and b.l[x] == b.l[y] and b.jj[x] == b.jj[y] and b.tz[x] == b.tz[y]:
                synth._set_one_body(b, Hrot, x, y, tr[x, y], 1)
    Hrot[0] = 0.0
    no = ob.normal_order(Hrot)
    # sqop_to_HF of the vacuum Hamiltonian is exactly this rotated, normal
    # ordered operator (its one-body rotation is the oracle's, not numpy's)
    sq = ob.sqop_to_hf(H, D)
    np.testing.assert_allclose(sq, no, rtol=1e-12, atol=1e-12)
    assert abs(no[0] - Hhf[0]) < 1e-5 * abs(Hhf[0])
    one = slice(1, b.off_fpp + b.npart * b.npart)
    assert np.max(np.abs(no[one] - Hhf[one])) < 1e-4
    np.testing.assert_array_equal(no[b.off_fpp + b.npart * b.npart:], Hhf[b.off_fpp + b.npart * b.npart:])


@pytest.mark.parametrize("emax", [2, 4])
def test_hf_brillouin_and_variational(emax):
    from eom_imsrg import synth
    b, ob = basis(emax), oracle_basis(emax)
    H = synth.hamiltonian_vacuum(b, seed=4, hw=20.0)
    Hhf, D, it = ob.hartree_fock(H)
    Hho = ob.normal_order(H)
    fph = slice(b.off_fph, b.off_fpp)
    assert np.max(np.abs(Hho[fph])) > 0.1
    assert np.max(np.abs(Hhf[fph])) < 1e-4
    assert Hhf[0] <= Hho[0] + 1e-9


def test_hf_errors_without_gpu():
    from eom_imsrg import _lib
    import ctypes as C
    b = basis(2)
    x = C.c_void_p(16)   # never dereferenced: the checks fail first
    assert _lib.lib.imsrg_hartree_fock(b.handle, None, 50, None, None, None, None) == _lib.IMSRG_EINVAL
    assert _lib.lib.imsrg_hartree_fock(b.handle, x, 0, None, None, None, None) == _lib.IMSRG_EINVAL
    assert _lib.lib.imsrg_hartree_fock(None, x, 50, None, None, None, None) == _lib.IMSRG_EINVAL
    n = _lib.lib.imsrg_basis_size(b.handle, _lib.SIZE["SPMAT"])
    assert n == int(oracle_basis(2).sp_blocks()[1][-1])


@pytest.mark.gpu
@pytest.mark.parametrize("emax", [4, 6])
def test_hf_matches_oracle(gpu, emax):
    """library HF (host SCF with Jacobi eigenvectors, two-body transform as
    two grouped GEMM families on the device) vs the oracle (LAPACK eigh, the
    reference's eight products): same iterations, E0, Fock matrix, two-body
    part and HF orbits to 1e-10"""
    from conftest import assert_rel
    from eom_imsrg import SqOp, interaction, synth
    b, ob = basis(emax), oracle_basis(emax)
    print(f"[hf] emax={emax}", flush=True)
    H = synth.hamiltonian_vacuum(b, seed=4, hw=20.0)
    ref, Dref, itref = ob.hartree_fock(H)
    Hd = SqOp.from_numpy(b, H, 1)
    e, it, D = interaction.hartree_fock(Hd)
    assert it == itref
    assert abs(e - ref[0]) < 1e-10 * abs(ref[0])
    assert_rel(Hd.numpy(), ref, rtol=1e-10, what=f"HF H emax {emax}")
    np.testing.assert_allclose(D, Dref, rtol=1e-10, atol=1e-12)


@pytest.mark.gpu
def test_flow_from_hf(gpu):
    """magnus_decouple from the HF-normal-ordered Hamiltonian: equal steps and
    E0 per step within 1e-8 MeV of the oracle's flow"""
    from eom_imsrg import Flow, SqOp, interaction, synth
    REF = float(np.float32(1e-6))
    b, ob = basis(4), oracle_basis(4)
    print("[hf flow] emax=4", flush=True)
    H = synth.hamiltonian_vacuum(b, seed=4, hw=20.0)
    ref, _, _ = ob.hartree_fock(H)
    Hd = SqOp.from_numpy(b, H, 1)
    interaction.hartree_fock(Hd)
    HSr, _, logr, str_ = ob.magnus_decouple(ref, 0.5, REF, REF, 0, REF)
    log = Flow(b, graphs=True).magnus_decouple(Hd, SqOp.zeros(b, -1))
    assert len(log) == str_
    assert np.array_equal(log[:, 3], logr[:, 3])
    assert np.max(np.abs(log[:, 1] - logr[:, 1])) < 1e-8


@pytest.mark.gpu
def test_sqop_to_hf_matches_oracle(gpu):
    """an observable (a second hermitian scalar operator) carried to the HF
    basis of H: library vs oracle to 1e-10"""
    from conftest import assert_rel
    from eom_imsrg import SqOp, interaction, synth
    b, ob = basis(4), oracle_basis(4)
    print("[sqop_to_hf] emax=4", flush=True)
    H = synth.hamiltonian_vacuum(b, seed=4, hw=20.0)
    Ob = synth.hamiltonian_vacuum(b, seed=9, hw=7.0)
    _, D, _ = ob.hartree_fock(H)
    ref = ob.sqop_to_hf(Ob, D)
    Hd = SqOp.from_numpy(b, H, 1)
    _, _, Dg = interaction.hartree_fock(Hd)
    Od = SqOp.from_numpy(b, Ob, 1)
    interaction.sqop_to_hf(Od, Dg)
    assert_rel(Od.numpy(), ref, rtol=1e-10, what="sqop_to_HF")
    with pytest.raises(ValueError):
        interaction.sqop_to_hf(SqOp.from_numpy(b, Ob, 1), Dg[:-1])   # coefs of the wrong length
