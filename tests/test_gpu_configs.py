"""The benchmark configurations at their own sizes on the GPU, against the
CPU oracle (run live where it takes seconds, else the committed fixtures
that tests/golden/make_configs.py wrote from it):

* eMax 8 (90 orbits): one [Omega, H] per element on the grouped and the
  per-block path; the per-block flow to |mbpt2| < 1e-6 (E0 and steps of the
  oracle's flow); the grouped hipGraph flow until |dE0| < 1e-8 MeV (E0 and
  commutators per step).
* eMax 10 (132 orbits): one [Omega, H] per element on the grouped path; the
  first Magnus step with BCH terms to 1e-10 (exit index, E0, mbpt2, seeded
  elements of H(s)) and the second step's MAGNUS_EXPAND (commutator count,
  norm of dOmega).
* 64 instances at eMax 6 as one batch: rows equal single-instance flows
  for a sample, E0 and steps equal the oracle's flows of instances 0..5
  (hw 16, 20 and 24).
* closed forms on the GPU path: mat_frob_norm of an all-ones operator and
  of single elements per block class; the flow of a Hamiltonian without a
  two-body part reaches sum (2j+1) * lowest eigenvalues.
"""
import json
import os

import numpy as np
import pytest
import torch

from conftest import assert_rel, basis, oracle_basis

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = float(np.float32(1e-6))
_CACHE = {}


def _fixture(name):
    return json.load(open(os.path.join(GOLDEN, name)))


def _h_and_omega(emax):
    """normal-ordered H (seed 0, hw 20) and Omega ~ N(0, 0.05) (seed 1)"""
    if emax not in _CACHE:
        from eom_imsrg import synth
        b, o = basis(emax), oracle_basis(emax)
        H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
        G = synth.antihermitian(b, seed=1, sigma=0.05)
        _CACHE[emax] = (H, G, None)
    return _CACHE[emax]


def _oracle_commutator(emax):
    H, G, C = _h_and_omega(emax)
    if C is None:
        C = oracle_basis(emax).commutator(G, -1, H, 1, 1)
        _CACHE[emax] = (H, G, C)
    return H, G, C


@pytest.mark.parametrize("path", [0, 1], ids=["grouped", "plain"])
def test_commutator_emax8_per_element(gpu, path):
    from eom_imsrg import SqOp, commutators as cm
    b = basis(8)
    print(f"[configs] emax 8 [Omega,H] path={path}", flush=True)
    H, G, ref = _oracle_commutator(8)
    b.set_path(path)
    try:
        RES = SqOp.zeros(b, 1)
        cm.commutator(SqOp.from_numpy(b, G, -1), SqOp.from_numpy(b, H, 1), RES)
        torch.cuda.synchronize()
        assert_rel(RES.numpy(), ref, what=f"[Omega,H] eMax 8 path {path}")
    finally:
        b.set_path(0)


def test_plain_flow_emax8_matches_oracle(gpu):
    """the per-block path's flow to |mbpt2| < 1e-6: steps and E0 of the
    oracle's flow (grad_emax8.json holds them for the same Hamiltonian)"""
    from eom_imsrg import Flow, SqOp
    fx = _fixture("grad_emax8.json")
    b = basis(8)
    print("[configs] emax 8 per-block flow", flush=True)
    H, _, _ = _h_and_omega(8)
    b.set_path(1)
    try:
        f = Flow(b, graphs=False)
        HS, G = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
        log = f.magnus_decouple(HS, G)
        torch.cuda.synchronize()
    finally:
        b.set_path(0)
    assert len(log) == fx["steps"], (len(log), fx["steps"])
    assert abs(log[-1, 1] - fx["E0"]) < 1e-8, (log[-1, 1], fx["E0"])


def test_grouped_flow_emax8_to_de0_1e8(gpu):
    """grouped + hipGraph flow until |E0(s) - E0(s - ds)| < 1e-8 MeV: every
    step's E0 within 1e-8 MeV and its commutator count equal to the oracle's"""
    from eom_imsrg import Flow, SqOp
    fx = _fixture("configs_emax8_flow.json")
    b = basis(8)
    print("[configs] emax 8 grouped flow, |dE0| < 1e-8", flush=True)
    H, _, _ = _h_and_omega(8)
    f = Flow(b, graphs=True)
    HS, G = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
    log = f.magnus_decouple(HS, G, crit_mode=1, crit_tol=1e-8, maxsteps=100)
    torch.cuda.synchronize()
    assert len(log) == fx["steps"], (len(log), fx["steps"])
    assert np.abs(log[:, 1] - np.array(fx["E0"])).max() < 1e-8
    assert [int(x) for x in log[:, 3]] == fx["commutators"]


def test_commutator_emax10_per_element(gpu):
    from eom_imsrg import SqOp, commutators as cm
    b = basis(10)
    print("[configs] emax 10 [Omega,H] (oracle ~10 s)", flush=True)
    H, G, ref = _oracle_commutator(10)
    RES = SqOp.zeros(b, 1)
    cm.commutator(SqOp.from_numpy(b, G, -1), SqOp.from_numpy(b, H, 1), RES)
    torch.cuda.synchronize()
    assert_rel(RES.numpy(), ref, what="[Omega,H] eMax 10")
    _CACHE.pop(10, None)


def test_first_magnus_steps_emax10(gpu):
    from eom_imsrg import Flow, SqOp, commutators as cm
    fx = _fixture("configs_emax10.json")
    b = basis(10)
    print("[configs] emax 10 first Magnus step, BCH to 1e-10", flush=True)
    from eom_imsrg import synth
    H = oracle_basis(10).normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    assert abs(H[0] - fx["E0_bare"]) < 1e-9
    f = Flow(b, graphs=True)
    Hd = SqOp.from_numpy(b, H, 1)
    HS, G, DG = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1), SqOp.zeros(b, -1)
    cm.build_gs_white(Hd, DG)
    e0, m, nm, nb = f.magnus_step(HS, G, DG, Hd, ds=0.5, bch_conv=1e-10, mag_conv=1e-10)
    torch.cuda.synchronize()
    assert nm == 0 and nb == fx["bch_index"], (nm, nb, fx["bch_index"])
    assert abs(e0 - fx["E0"]) < 1e-8 and abs(m - fx["mbpt2"]) <= 1e-10 * abs(fx["mbpt2"])
    idx = np.asarray(fx["index"], np.int64)
    got = HS.data[torch.from_numpy(idx).cuda()].cpu().numpy()
    assert_rel(got, np.asarray(fx["HS"]), what="H(s=0.5) sample")
    assert abs(cm.mat_frob_norm(HS) - fx["norm"]) <= 1e-10 * fx["norm"]
    nm2 = f.MAGNUS_EXPAND(DG, G, conv=1e-10)
    torch.cuda.synchronize()
    assert nm2 == fx["magnus2_commutators"]
    assert abs(cm.mat_frob_norm(DG) - fx["magnus2_norm"]) <= 1e-10 * fx["magnus2_norm"]


def test_batch_64_emax6(gpu):
    from eom_imsrg import Batch, Flow, SqOp, synth
    fx = _fixture("configs_batch_emax6.json")
    b, o = basis(6), oracle_basis(6)
    ni = 64
    print(f"[configs] batch {ni} x emax 6", flush=True)
    Hs = []
    for k in range(ni):
        seed, scale, hw = synth.instance_parameters(k)
        Hs.append(o.normal_order(synth.hamiltonian_vacuum(b, seed=seed, hw=hw, v_scale=scale)))
    H = np.stack(Hs)
    bt = Batch(b, ni, graphs=True)
    HS = torch.from_numpy(H).cuda()
    G = bt.zeros()
    e0, steps, nmax = bt.magnus_decouple(HS, G)
    torch.cuda.synchronize()
    assert nmax == steps.max()
    for row in fx["instances"]:
        k = row["instance"]
        assert steps[k] == row["steps"], (k, steps[k], row["steps"])
        assert abs(e0[k] - row["E0"]) < 1e-8, (k, e0[k], row["E0"])
    assert {row["hw"] for row in fx["instances"]} == {16.0, 20.0, 24.0}
    f = Flow(b, graphs=True)
    for k in (0, 21, 42, 63):
        hs, g = SqOp.from_numpy(b, H[k], 1), SqOp.zeros(b, -1)
        log = f.magnus_decouple(hs, g)
        torch.cuda.synchronize()
        assert len(log) == steps[k] and log[-1, 1] == e0[k], (k, len(log), steps[k], log[-1, 1], e0[k])
        assert_rel(HS[k].cpu().numpy(), hs.numpy(), rtol=1e-14, what=f"batch row {k}")


# ------------------------------------------------------------ closed forms on the GPU path
def test_frob_norm_closed_forms_gpu(gpu):
    from eom_imsrg import SqOp, commutators as cm
    b = basis(4)
    n_q = b.dims.sum(axis=1).astype(np.float64)
    ref = np.sqrt(float(b.norb) ** 2 + float(np.sum(n_q * n_q)))
    assert abs(cm.mat_frob_norm(SqOp.from_numpy(b, np.ones(b.oplen), 1)) - ref) < 1e-12 * ref
    A, P = b.nhole, b.npart
    weights = (1.0, 2.0, 2.0, 1.0, 1.0, 2.0)
    cases = [(b.off_fhh, 1.0), (b.off_fph, 2.0), (b.off_fpp, 1.0)]
    for g in range(6):
        q = next(q for q in range(b.nblocks) if np.prod(b.block_shape(q, g)) > 0)
        cases.append((int(b.goff[q, g]), weights[g]))
    for off, w in cases:
        op = np.zeros(b.oplen)
        op[off] = -3.25
        assert abs(cm.mat_frob_norm(SqOp.from_numpy(b, op, 1)) - 3.25 * np.sqrt(w)) < 1e-14, (off, w)


def test_one_body_flow_reaches_exact_energy_gpu(gpu):
    from eom_imsrg import Flow, SqOp, synth
    from jscheme import ljt_blocks
    b, o = basis(4), oracle_basis(4)
    print("[configs] one-body flow closed form", flush=True)
    Hv = synth.hamiltonian_vacuum(b, seed=3, hw=20.0)
    Hv[b.off_fpp + b.npart * b.npart:] = 0.0
    H = o.normal_order(Hv)
    A, P = b.nhole, b.npart
    F = np.zeros((b.norb, b.norb))
    F[np.ix_(b.holes, b.holes)] = H[b.off_fhh: b.off_fhh + A * A].reshape(A, A, order="F")
    F[np.ix_(b.parts, b.parts)] = H[b.off_fpp: b.off_fpp + P * P].reshape(P, P, order="F")
    fph = H[b.off_fph: b.off_fph + P * A].reshape(P, A, order="F")
    F[np.ix_(b.parts, b.holes)] = fph
    F[np.ix_(b.holes, b.parts)] = fph.T
    exact = 0.0
    for blk in ljt_blocks(b):
        nh = int(sum(b.hole[a] for a in blk))
        if nh:
            exact += (b.jj[blk[0]] + 1) * float(np.sum(np.linalg.eigvalsh(F[np.ix_(blk, blk)])[:nh]))
    f = Flow(b, graphs=True)
    HS, G = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
    log = f.magnus_decouple(HS, G, bch_conv=1e-13, crit_mode=1, crit_tol=1e-12, maxsteps=200)
    torch.cuda.synchronize()
    assert 0 < len(log) < 200
    assert abs(HS.E0 - exact) < 1e-11 * abs(exact), (HS.E0, exact)
