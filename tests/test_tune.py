"""GEMM tile autotuner: the stored choice is keyed by device and basis,
read back from the cache file (or the shipped MI355X table), and applied
through imsrg_set_gemm_tile; on the GPU the measured choice must leave the
commutator unchanged to rounding (XCD packing and the LDS-staged shapes
reorder tiles, not sums; the register-direct shapes also regroup each
K sum, which moves the last bits)."""
import json

import numpy as np
import pytest
import torch

from conftest import assert_rel, basis


def test_choice_round_trip(tmp_path, monkeypatch):
    from eom_imsrg import tune
    monkeypatch.setenv("EOM_IMSRG_TUNE_FILE", str(tmp_path / "tiles.json"))
    b = basis(2)
    before = tune.current(b)
    assert tune.lookup(b) is None or isinstance(tune.lookup(b), dict)
    choice = {"PPPP": 1, "PPX": 8, "XX": 2, "PH": 10}
    (tmp_path / "tiles.json").write_text(json.dumps({tune.basis_key(b): choice}))
    assert tune.lookup(b) == choice
    got = tune.tune(b)                 # from the cache: applied, not measured
    assert got == {"choice": choice, "source": "cache", "key": tune.basis_key(b)}
    assert tune.current(b) == choice
    tune.apply(b, before)
    assert tune.current(b) == before


def test_basis_key_separates_bases():
    from eom_imsrg import tune
    assert tune.basis_key(basis(2)) != tune.basis_key(basis(4))
    assert tune.basis_key(basis(2)) == tune.basis_key(basis(2))


def test_shipped_table_is_valid_json():
    from eom_imsrg import tune
    d = json.load(open(tune.SHIPPED))
    for k, c in d.items():
        assert set(c) == set(tune.FAMILIES), k
        assert all(v in tune.CANDIDATES for v in c.values()), k


@pytest.mark.gpu
def test_measured_choice_keeps_results(gpu, tmp_path, monkeypatch):
    from eom_imsrg import SqOp, commutators as cm, synth, tune
    monkeypatch.setenv("EOM_IMSRG_TUNE_FILE", str(tmp_path / "tiles.json"))
    b = basis(4)
    print("[tune] emax=4", flush=True)
    before = tune.current(b)
    L = SqOp.from_numpy(b, synth.antihermitian(b, 3, 0.1), -1)
    R = SqOp.from_numpy(b, synth.hermitian_random(b, 4, 1.0), 1)
    ref = SqOp.zeros(b, 1)
    cm.commutator(L, R, ref)
    torch.cuda.synchronize()
    out = tune.tune(b, iters=2, force=True)
    assert out["source"] == "measured" and set(out["choice"]) == set(tune.FAMILIES)
    assert json.load(open(tmp_path / "tiles.json"))[tune.basis_key(b)] == out["choice"]
    got = SqOp.zeros(b, 1)
    cm.commutator(L, R, got)
    torch.cuda.synchronize()
    assert_rel(got.numpy(), ref.numpy(), rtol=1e-12, what="commutator after tuning")
    tune.apply(b, before)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5, 6, 7])
def test_every_tile_shape_matches_oracle(gpu, tile):
    """every GEMM family on one tile shape (64x64, 128x64, 128x128, 64x32,
    32x64, 32x32 through LDS, 128x64 and 64x64 register-direct, with XCD
    packing on odd shapes): the commutator equals the oracle's, including
    eMax 6 blocks larger than the tiles"""
    from eom_imsrg import SqOp, commutators as cm, synth, tune
    from conftest import oracle_basis
    b, o = basis(6), oracle_basis(6)
    print(f"[tile] shape={tile}", flush=True)
    before = tune.current(b)
    L = synth.antihermitian(b, 3, 0.1)
    R = synth.hermitian_random(b, 4, 1.0)
    ref = o.commutator(L, -1, R, 1, 1)
    tune.apply(b, {fam: tile | (8 if tile % 2 else 0) for fam in tune.FAMILIES})
    try:
        got = SqOp.zeros(b, 1)
        cm.commutator(SqOp.from_numpy(b, L, -1), SqOp.from_numpy(b, R, 1), got)
        torch.cuda.synchronize()
        assert_rel(got.numpy(), ref, what=f"[L,R] with tile shape {tile}")
    finally:
        tune.apply(b, before)
