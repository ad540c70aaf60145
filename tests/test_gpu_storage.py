"""Host <-> device operator components through imsrg_op_put / imsrg_op_get
(the per-block mirror a Fortran sq_op caller uses) and imsrg_alloc."""
import ctypes as C

import numpy as np
import pytest

from conftest import basis

pytestmark = pytest.mark.gpu


def test_op_put_get_round_trip(gpu):
    from eom_imsrg import _lib, synth
    b = basis(4)
    print("[storage] emax=4", flush=True)
    lib = _lib.lib
    ref = synth.hermitian_random(b, 5, 1.0)
    ref[0] = 3.25
    dev = C.c_void_p()
    _lib.check(lib.imsrg_alloc(b.oplen, C.byref(dev)))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    A, P = b.nhole, b.npart
    comps = [(-1, 0, ref[0:1]), (-1, 1, ref[b.off_fhh:b.off_fhh + A * A]),
             (-1, 2, ref[b.off_fph:b.off_fph + P * A]), (-1, 3, ref[b.off_fpp:b.off_fpp + P * P])]
    for q in range(b.nblocks):
        for g in range(6):
            n1, n2 = b.block_shape(q, g)
            o = int(b.goff[q, g])
            comps.append((q, g, ref[o:o + n1 * n2]))
    for q, g, x in comps:
        x = np.ascontiguousarray(x)
        _lib.check(lib.imsrg_op_put(b.handle, dev, q, g, dp(x)))
    flat = np.zeros(b.oplen)
    _lib.check(lib.imsrg_op_get(b.handle, dev, -1, 0, dp(flat[0:1])))
    assert flat[0] == 3.25
    for q, g, x in comps[1:]:
        out = np.zeros(len(x))
        _lib.check(lib.imsrg_op_get(b.handle, dev, q, g, dp(out)))
        assert np.array_equal(out, x), (q, g)
    assert lib.imsrg_op_get(b.handle, dev, b.nblocks, 0, dp(flat)) == _lib.IMSRG_EINVAL
    assert lib.imsrg_op_put(b.handle, dev, -1, 4, dp(flat)) == _lib.IMSRG_EINVAL
    _lib.check(lib.imsrg_free(dev))
