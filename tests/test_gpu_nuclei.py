"""Other reference states than 16O: the plans (pair lists, 122 partners,
222_ph chunks, cross-coupled blocks) depend on which orbits are holes, so
the full commutator (grouped and per-block paths) and a flow are checked
against the oracle for

* 4He   (0s1/2 protons and neutrons),
* 40Ca  (every orbit with 2n + l <= 2),
* 22O   (16O plus the 0d5/2 neutrons: proton and neutron holes differ, so
         channels of opposite Tz have different shapes),
* 14C   (16O without the 0p1/2 protons: the proton 0p1/2 orbit is a
         particle, its spin-orbit partner a hole).
"""
import numpy as np
import pytest
import torch

from conftest import assert_rel

pytestmark = pytest.mark.gpu

REF = float(np.float32(1e-6))


def _holes(o, name):
    n, l, jj, tz = (np.asarray(o[k]) for k in ("n", "l", "jj", "tz"))
    if name == "He4":
        h = (n == 0) & (l == 0)
    elif name == "Ca40":
        h = 2 * n + l <= 2
    elif name == "O22":
        h = ((n == 0) & (l <= 1)) | ((n == 0) & (l == 2) & (jj == 5) & (tz == 1))
    elif name == "C14":
        h = ((n == 0) & (l <= 1)) & ~((l == 1) & (jj == 1) & (tz == -1))
    else:
        raise ValueError(name)
    return h.astype(np.int32)


_CACHE = {}


def _bases(name, emax):
    key = (name, emax)
    if key not in _CACHE:
        from eom_imsrg import Basis, hk_orbits
        from oracle.oracle import OracleBasis
        o = hk_orbits(emax)
        hole = _holes(o, name)
        b = Basis(o["n"], o["l"], o["jj"], o["tz"], hole, o["e"])
        _CACHE[key] = (b, OracleBasis(b.n, b.l, b.jj, b.tz, b.hole))
    return _CACHE[key]


@pytest.mark.parametrize("name", ["He4", "Ca40", "O22", "C14"])
@pytest.mark.parametrize("path", [0, 1], ids=["grouped", "plain"])
def test_commutator_vs_oracle(gpu, name, path):
    from eom_imsrg import SqOp, commutators as cm, synth
    b, o = _bases(name, 4)
    print(f"[nucleus] {name} emax=4 path={path} holes={int(b.hole.sum())}", flush=True)
    b.set_path(path)
    try:
        H = o.normal_order(synth.hamiltonian_vacuum(b, seed=3, hw=20.0))
        G = synth.antihermitian(b, seed=4, sigma=0.05)
        for (L, Lh, R, Rh) in ((G, -1, H, 1), (G, -1, synth.antihermitian(b, seed=5, sigma=0.3), -1)):
            RES = SqOp.zeros(b, -Lh * Rh)
            cm.commutator(SqOp.from_numpy(b, L, Lh), SqOp.from_numpy(b, R, Rh), RES)
            torch.cuda.synchronize()
            assert_rel(RES.numpy(), o.commutator(L, Lh, R, Rh, -Lh * Rh), what=f"{name} [L,R] Lh={Lh} Rh={Rh}")
    finally:
        b.set_path(0)


@pytest.mark.parametrize("name", ["O22", "C14"])
def test_flow_vs_oracle(gpu, name):
    """magnus_decouple at eMax 4: equal steps and commutator counts, E0 per
    step within 1e-8 MeV"""
    from eom_imsrg import Flow, SqOp, synth
    b, o = _bases(name, 4)
    print(f"[nucleus flow] {name} emax=4", flush=True)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    HSr, _, logr, str_ = o.magnus_decouple(H, 0.5, REF, REF, 0, REF)
    f = Flow(b, graphs=True)
    HS = SqOp.from_numpy(b, H, 1)
    log = f.magnus_decouple(HS, SqOp.zeros(b, -1))
    torch.cuda.synchronize()
    assert len(log) == str_
    assert np.array_equal(log[:, 3], logr[:, 3])
    assert np.max(np.abs(log[:, 1] - logr[:, 1])) < 1e-8
    assert abs(HS.E0 - HSr[0]) < 1e-8


def test_flow_gradient_vs_oracle_c14(gpu):
    """dE0/dT of the whole 14C flow (Omega checkpointed every 4 steps)
    against the oracle's reverse sweep at eMax 4, 1e-8 relative"""
    from eom_imsrg import Flow, adjoint, synth
    b, o = _bases("C14", 4)
    print("[nucleus flow_grad] C14 emax=4", flush=True)
    T = synth.hamiltonian_vacuum(b, 0)
    e0_o, Tb_o, st_o = o.flow_grad(T, 0.5, REF, REF, REF, ckpt=4)
    e0, Tb, st = adjoint.flow_grad(Flow(b, graphs=True), torch.from_numpy(T).cuda(), ckpt=4)
    assert st == st_o and abs(e0 - e0_o) < 1e-8
    assert_rel(adjoint.symmetric_gradient(b, Tb.cpu().numpy(), 1), adjoint.symmetric_gradient(b, Tb_o, 1), rtol=1e-8,
               what="dE0/dT 14C")


@pytest.mark.parametrize("name", ["He4", "O22", "C14"])
def test_hartree_fock_vs_oracle(gpu, name):
    """Hartree-Fock (calc_HF) for other reference states at eMax 4: the
    occupied HF orbits of a block are its lowest eigenvectors, as many as
    the block has hole orbits (22O fills the neutron 0d5/2 block's lower
    orbit, 14C leaves the proton 0p1/2 block empty); same iterations, E0,
    Fock matrix, two-body part and orbits as the oracle to 1e-10"""
    from eom_imsrg import SqOp, interaction, synth
    b, o = _bases(name, 4)
    print(f"[nucleus hf] {name} emax=4", flush=True)
    H = synth.hamiltonian_vacuum(b, seed=4, hw=20.0)
    ref, Dref, itref = o.hartree_fock(H)
    Hd = SqOp.from_numpy(b, H, 1)
    e, it, D = interaction.hartree_fock(Hd)
    assert it == itref
    assert abs(e - ref[0]) < 1e-10 * abs(ref[0])
    assert_rel(Hd.numpy(), ref, rtol=1e-10, what=f"HF {name}")
    np.testing.assert_allclose(D, Dref, rtol=1e-10, atol=1e-12)
