"""The operator stream of vectorize / repackage (basic_IMSRG.f90:3929-4053)
and the .bin file of write_omega_checkpoint / read_omega_checkpoint
(3215-3407).

The reference holds no .bin file to compare against, so the oracle's
restatement is pinned by what the loop nest fixes: the stream length against
read_omega_checkpoint's own neq count, the visiting order read off an
operator whose elements are their own flat indices, and round trips.  The
GPU kernels must then equal the oracle bit for bit (it is a permutation),
and the file must be the stream as raw little-endian doubles."""
import os

import numpy as np
import pytest
import torch

from conftest import basis, oracle_basis


def _neq_reference_count(b):
    """neq as read_omega_checkpoint counts it (basic_IMSRG.f90:3338-3348)"""
    A, N = b.nhole, b.norb
    neq = 1 + A * A + A * (N - A) + (N - A) ** 2
    for q in range(b.nblocks):
        npp, nph, nhh = (int(x) for x in b.dims[q])
        neq += (nhh * nhh + nhh + nph * nph + nph + npp * npp + npp) // 2 + nph * npp + nhh * npp + nhh * nph
    return neq


@pytest.mark.parametrize("emax", [2, 4, 6])
def test_stream_length(emax):
    b, o = basis(emax), oracle_basis(emax)
    n = _neq_reference_count(b)
    assert b.neq == n
    assert len(o.vectorize(np.zeros(o.oplen))) == n


def test_stream_visiting_order():
    """with op[x] = x the stream lists flat indices: E0, fhh, fpp, fph
    column-major, then per channel gam 1..6; square blocks column by column
    from the diagonal down, rectangular ones whole, column-major"""
    b, o = basis(4), oracle_basis(4)
    op = np.arange(o.oplen, dtype=np.float64)
    v = o.vectorize(op).astype(np.int64)
    A, P = b.nhole, b.npart
    exp = [0]
    exp += [b.off_fhh + r + c * A for c in range(A) for r in range(A)]
    exp += [b.off_fpp + r + c * P for c in range(P) for r in range(P)]
    exp += [b.off_fph + r + c * P for c in range(A) for r in range(P)]
    for q in range(b.nblocks):
        for g in range(6):
            nr, nc = b.block_shape(q, g)
            off = int(b.goff[q, g])
            for c in range(nc):
                for r in range(c if g in (0, 3, 4) else 0, nr):
                    exp.append(off + r + c * nr)
    assert np.array_equal(v, np.array(exp))
    # every stored element of a rectangular block and the lower triangle of
    # a square one appears once; nothing from a square upper triangle
    assert len(set(v.tolist())) == len(v)


def test_round_trips():
    """repackage(vectorize(H), herm) = H for hermitian and antihermitian
    operators; a square diagonal comes back as herm * v"""
    from eom_imsrg import synth
    b, o = basis(4), oracle_basis(4)
    H = o.normal_order(synth.hamiltonian_vacuum(b, 0))
    assert np.array_equal(o.repackage(o.vectorize(H), 1), H)
    G = synth.antihermitian(b, 1, 0.05)
    assert np.array_equal(o.repackage(o.vectorize(G), -1), G)
    v = o.vectorize(H)
    back = o.repackage(v, -1)
    q = next(q for q in range(b.nblocks) if b.dims[q, 0] > 1)
    n = int(b.dims[q, 0])
    X = back[int(b.goff[q, 0]): int(b.goff[q, 0]) + n * n].reshape(n, n)
    Xh = H[int(b.goff[q, 0]): int(b.goff[q, 0]) + n * n].reshape(n, n)
    assert np.array_equal(np.diag(X), -np.diag(Xh))
    assert np.array_equal(np.tril(X.T, -1), np.tril(Xh.T, -1))          # lower (row > col) kept
    assert np.array_equal(np.triu(X.T, 1), -np.tril(Xh.T, -1).T)          # upper = herm * lower


@pytest.mark.gpu
def test_gpu_vectorize_repackage_and_file(gpu, tmp_path):
    from eom_imsrg import SqOp, synth
    for emax in (4, 6):
        b, o = basis(emax), oracle_basis(emax)
        print(f"[op file] emax={emax}", flush=True)
        H = o.normal_order(synth.hamiltonian_vacuum(b, 3))
        G = synth.antihermitian(b, 4, 0.05)
        for arr, herm in ((H, 1), (G, -1)):
            op = SqOp.from_numpy(b, arr, herm)
            v = op.vectorize()
            torch.cuda.synchronize()
            ref = o.vectorize(arr)
            assert np.array_equal(v.cpu().numpy(), ref)
            # repackage from a stream whose upper partners disagree with herm
            w = torch.from_numpy(ref * 1.5).cuda()
            back = SqOp.repackage(b, w, -herm)
            torch.cuda.synchronize()
            assert np.array_equal(back.numpy(), o.repackage(ref * 1.5, -herm))
            path = os.path.join(str(tmp_path), f"op_e{emax}_{herm}.bin")
            op.write_bin(path)
            raw = np.fromfile(path, dtype="<f8")
            assert raw.size == b.neq and np.array_equal(raw, ref)
            rd = SqOp.read_bin(b, path, herm)
            assert np.array_equal(rd.numpy(), arr)
    # a file of another basis is refused
    from eom_imsrg import _lib
    with pytest.raises(_lib.ImsrgError):
        SqOp.read_bin(basis(4), os.path.join(str(tmp_path), "op_e6_1.bin"))


def test_omega_checkpoint_names():
    """write_omega_checkpoint's names (basic_IMSRG.f90:3239-3249): s as F6.3,
    left adjusted, a '0' appended below 10, asterisks when it does not fit"""
    from eom_imsrg import SqOp
    name = SqOp.omega_checkpoint_name
    assert name("x/o16_hw20", 0.5) == "x/o16_hw20_s0.5000.bin"
    assert name("p", 5.5) == "p_s5.5000.bin"
    assert name("p", 9.5) == "p_s9.5000.bin"
    assert name("p", 10.0) == "p_s10.000.bin"
    assert name("p", 95.5) == "p_s95.500.bin"
    assert name("p", -0.5) == "p_s-0.5000.bin"
    assert name("p", 100.0) == "p_s******.bin"
