"""magnus_decouple with the reference's Omega checkpoints
(IMSRG_MAGNUS.f90:64-110, write_omega_checkpoint / read_omega_checkpoint,
basic_IMSRG.f90:3215-3407) on the GPU: which files a flow writes and
removes, that a file holds Omega after the steps before it, and that a flow
restarted from the newest file continues the uninterrupted flow (same s,
steps and E0 per step)."""
import os

import numpy as np
import pytest
import torch

from conftest import basis, oracle_basis

pytestmark = pytest.mark.gpu


def test_checkpoint_files_and_restart(gpu, tmp_path):
    from eom_imsrg import Flow, SqOp, synth, _lib
    b, o = basis(4), oracle_basis(4)
    print("[checkpoint] emax=4 chk=4", flush=True)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    prefix = str(tmp_path / "o16_eMax04_hw20")
    f = Flow(b, graphs=True)
    HS, G = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
    log, st, sx = f.magnus_decouple_chk(HS, G, prefix, chk=4, crit_mode=1, crit_tol=1e-12)
    torch.cuda.synchronize()
    assert sx == -1.0 and st == len(log) and st > 28, (sx, st, len(log))
    # written before steps 4, 8, ...: Omega after 3, 7, ... steps (s = 1.5, 3.5, ...);
    # writing s removes the file of s - 12
    written = [0.5 * (k - 1) for k in range(4, st + 1, 4)]
    kept = [s for s in written if s + 12.0 not in written]
    assert len(kept) < len(written)
    names = {SqOp.omega_checkpoint_name(prefix, s) for s in kept}
    assert {str(p) for p in tmp_path.iterdir()} == names
    # the newest file = the stream of Omega after its steps
    s_new = max(kept)
    k = int(round(s_new / 0.5))
    g2 = Flow(b, graphs=True)
    HS2, G2 = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
    with pytest.raises(_lib.ImsrgError):
        g2.magnus_decouple(HS2, G2, crit_mode=1, crit_tol=1e-12, maxsteps=k)   # stops after k steps
    vec = G2.vectorize().cpu().numpy()
    raw = np.fromfile(SqOp.omega_checkpoint_name(prefix, s_new), dtype=np.float64)
    assert np.array_equal(raw, vec)
    # restart from it
    HS3, G3 = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
    log3, st3, sx3 = f.magnus_decouple_chk(HS3, G3, prefix, chk=4, crit_mode=1, crit_tol=1e-12)
    torch.cuda.synchronize()
    assert sx3 == s_new and st3 == st and len(log3) == st - k, (sx3, st3, len(log3))
    assert np.abs(log3[:, 0] - log[k:, 0]).max() == 0.0
    assert np.abs(log3[:, 1] - log[k:, 1]).max() < 1e-10
    assert abs(HS3.E0 - HS.E0) < 1e-10


def test_checkpoint_default_interval_and_no_file(gpu, tmp_path):
    """chk = 0: the reference's 24 steps below eMax 14; a prefix with no file
    is a fresh start"""
    from eom_imsrg import Flow, SqOp, synth
    b, o = basis(4), oracle_basis(4)
    H = o.normal_order(synth.hamiltonian_vacuum(b, seed=0, hw=20.0))
    prefix = str(tmp_path / "run")
    f = Flow(b, graphs=True)
    HS, G = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
    log, st, sx = f.magnus_decouple_chk(HS, G, prefix, chk=0, crit_mode=1, crit_tol=1e-12)
    torch.cuda.synchronize()
    assert sx == -1.0 and st >= 24
    assert {str(p) for p in tmp_path.iterdir()} == {SqOp.omega_checkpoint_name(prefix, 11.5)}
    ref = Flow(b, graphs=True)
    HSr, Gr = SqOp.from_numpy(b, H, 1), SqOp.zeros(b, -1)
    logr = ref.magnus_decouple(HSr, Gr, crit_mode=1, crit_tol=1e-12)
    assert len(logr) == st and np.array_equal(logr, log)
