"""Seeded synthetic inputs: SplitMix64 known answers, Box-Muller moments, and
the structure of the generated operators (symmetry, Pauli zeros, herm)."""
import numpy as np

from conftest import basis
from jscheme import each_element, two_body_part


def test_splitmix64_known_answers():
    """published SplitMix64 outputs for seed 0 and seed 1234567"""
    from eom_imsrg.synth import SplitMix64
    assert [int(x) for x in SplitMix64(0).raw(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    r = SplitMix64(1234567)
    assert [int(x) for x in r.raw(2)] == [6457827717110365317, 3203168211198807973]
    # streaming in pieces gives the same sequence
    a = SplitMix64(99).raw(10)
    r = SplitMix64(99)
    b = np.concatenate([r.raw(3), r.raw(7)])
    assert (a == b).all()


def test_normals_moments():
    from eom_imsrg.synth import SplitMix64
    x = SplitMix64(5).normal(200_001)
    assert len(x) == 200_001
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    u = SplitMix64(5).uniform(1000)
    assert (u >= 0).all() and (u < 1).all()


def _blocks(b, op):
    for q in range(b.nblocks):
        for g in range(6):
            n1, n2 = b.block_shape(q, g)
            if n1 * n2:
                o = int(b.goff[q, g])
                yield q, g, op[o: o + n1 * n2].reshape(n1, n2, order="F")


def test_operator_structure():
    from eom_imsrg import synth
    b = basis(4)
    H = synth.hamiltonian_vacuum(b, 0)
    Om = synth.antihermitian(b, 1, 0.05)
    assert np.array_equal(H, synth.hamiltonian_vacuum(b, 0))
    for q, g, X in _blocks(b, H):
        if g in (0, 3, 4):
            assert np.array_equal(X, X.T)
    for q, g, X in _blocks(b, Om):
        if g in (0, 3, 4):
            assert np.array_equal(X, -X.T)
    A, P = b.nhole, b.npart
    fhh = H[b.off_fhh: b.off_fhh + A * A].reshape(A, A, order="F")
    assert np.array_equal(fhh, fhh.T)
    ohh = Om[b.off_fhh: b.off_fhh + A * A].reshape(A, A, order="F")
    assert np.array_equal(ohh, -ohh.T)
    # one-body diagonal = hw (2n + l + 3/2); couplings only inside (l, j, tz) blocks
    fpp = H[b.off_fpp: b.off_fpp + P * P].reshape(P, P, order="F")
    for k, a in enumerate(b.parts):
        assert fpp[k, k] == 20.0 * (b.e[a] + 1.5)
        for m, c in enumerate(b.parts):
            if (b.l[a], b.jj[a], b.tz[a]) != (b.l[c], b.jj[c], b.tz[c]):
                assert fpp[k, m] == 0.0
    # Pauli: pairs (a,a) carry zero at odd J
    for q, g, ix, jx, (a, bb), (c, d), J in each_element(b):
        if (J // 2) % 2 == 1 and (a == bb or c == d):
            n1, _ = b.block_shape(q, g)
            assert H[int(b.goff[q, g]) + ix + jx * n1] == 0.0
    # two-body damping exp(-(e_a+e_b+e_c+e_d)/8) keeps |V| bounded
    assert np.abs(H[two_body_part(b)]).max() < 6.0


def test_instance_parameters():
    from eom_imsrg import synth
    ps = [synth.instance_parameters(k) for k in range(64)]
    assert [p[0] for p in ps] == list(range(64))
    assert all(0.8 <= p[1] <= 1.2 for p in ps)
    assert {p[2] for p in ps} == {16.0, 20.0, 24.0}
    assert ps == [synth.instance_parameters(k) for k in range(64)]
