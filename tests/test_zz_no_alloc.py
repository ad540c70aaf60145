"""imsrg_commutator_222_ph_noalloc (the 222_ph term with its WCC workspace in
the library's fixed device array) against imsrg_commutator_222_ph with a
caller workspace, through the C-ABI: identical RES on two calls in a row,
the second on a smaller basis (the array keeps the first call's contents),
and the documented refusals (workspace larger than the array, basis without
device tables)."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import basis

pytestmark = pytest.mark.gpu


def _case(emax, seed):
    from eom_imsrg import _lib, CCMat, CCWork, SqOp, synth, commutators as cm
    b = basis(emax)
    print(f"[noalloc] emax={emax} wcc={b.wcclen} doubles", flush=True)
    L = SqOp.from_numpy(b, synth.antihermitian(b, seed, 1.0), -1)
    R = SqOp.from_numpy(b, synth.hermitian_random(b, seed + 1, 1.0), 1)
    prior = synth.hermitian_random(b, seed + 2, 0.3)
    LCC, RCC = CCMat(b, -1), CCMat(b, 1)
    cm.calculate_cross_coupled(L, LCC)
    cm.calculate_cross_coupled(R, RCC)
    RES_ref = SqOp.from_numpy(b, prior, 1)
    W = CCWork(b)
    _lib.check(_lib.lib.imsrg_commutator_222_ph(b.handle, LCC.ptr, -1, RCC.ptr, 1, RES_ref.ptr, 1, W.ptr, None))
    RES = SqOp.from_numpy(b, prior, 1)
    _lib.check(_lib.lib.imsrg_commutator_222_ph_noalloc(b.handle, LCC.ptr, -1, RCC.ptr, 1, RES.ptr, 1, None))
    torch.cuda.synchronize()
    got, ref = RES.numpy(), RES_ref.numpy()
    assert np.abs(ref - prior).max() > 1e-3      # the term did change RES
    assert np.array_equal(got, ref), float(np.abs(got - ref).max())


def test_noalloc_222_ph_matches_caller_workspace(gpu):
    _case(6, 31)
    _case(4, 41)     # second call, smaller basis: stale array contents must not leak


def test_noalloc_222_ph_refusals(gpu):
    from eom_imsrg import _lib, Basis
    p = C.c_void_p(16)
    big = basis(8)
    assert big.wcclen > (64 << 20) // 8
    rc = _lib.lib.imsrg_commutator_222_ph_noalloc(big.handle, p, -1, p, 1, p, 1, None)
    assert rc == _lib.IMSRG_EINVAL and "64 MiB" in _lib.last_error()
    fresh = Basis.oxygen16(2)      # no compute call yet: no device tables
    rc = _lib.lib.imsrg_commutator_222_ph_noalloc(fresh.handle, p, -1, p, 1, p, 1, None)
    assert rc == _lib.IMSRG_EINVAL and "device tables" in _lib.last_error()
