"""Algorithmic work of one scalar commutator [L, R] on a basis.

F_alg counts what the reference computes with dgemm: the 18 per-channel
GEMM shapes of commutator_222_pp_hh (commutators.f90:611-755), the
per-cross-coupled-block GEMM of commutator_222_ph (commutators.f90:799) and
the 8 one-body GEMMs of commutator_111 (commutators.f90:56-76); 2mnk flops
each.  B_alg = 24 (E1 + E2) bytes: read L, read R, write RES once.
Per-kernel figures are what each kernel algorithmically needs.
"""
from __future__ import annotations

from .basis import Basis


def _g(m, n, k):
    return 2.0 * m * n * k


def pphh_flops(b: Basis) -> float:
    f = 0.0
    for q in range(b.nblocks):
        np_, nb, nh = (int(x) for x in b.dims[q])
        f += _g(np_, np_, np_) + _g(nh, nh, nh) + _g(nh, nh, np_) + _g(np_, np_, nh)
        f += 2 * _g(np_, nh, np_) + 2 * _g(np_, nh, nh)
        f += _g(nb, nb, nh) + _g(nb, nb, np_)
        f += 2 * _g(nb, nh, nh) + 2 * _g(np_, nb, np_) + 2 * _g(nb, nh, np_) + 2 * _g(np_, nb, nh)
    return f


def pppp_flops(b: Basis) -> float:
    """the pppp tile-pair kernel: L1 R1 (np^3) and the rank-nh L3 R3^T part"""
    return sum(_g(int(d[0]), int(d[0]), int(d[0])) + _g(int(d[0]), int(d[0]), int(d[2])) for d in b.dims)


def ph_gemm_flops(b: Basis) -> float:
    return sum(2.0 * int(r) * int(r) * int(c) for r, c in zip(b.rlen, b.ncol))


def ph_gemm_bytes(b: Basis) -> float:
    return 8.0 * (b.wcclen + 2 * b.cclen)


def onebody_flops(b: Basis) -> float:
    A, P = b.nhole, b.npart
    return (_g(A, A, A) + _g(A, A, P) + _g(P, A, A) + _g(P, A, P) + _g(P, A, A) + _g(P, A, P)
            + _g(P, P, A) + _g(P, P, P))


def e1(b: Basis) -> int:
    return b.nhole * b.nhole + b.npart * b.nhole + b.npart * b.npart


def commutator_flops(b: Basis) -> float:
    return pphh_flops(b) + ph_gemm_flops(b) + onebody_flops(b)


def commutator_bytes(b: Basis) -> float:
    return 24.0 * (e1(b) + b.e2_count())


def scatter_bytes(b: Basis) -> float:
    """inverse Pandya: every WCC element read once, RES read and written"""
    return 8.0 * (b.wcclen + 2 * b.e2_count())


def pppp_bytes(b: Basis) -> float:
    """L1, R1 read, RES1 read + written, w1/w2 pppp written"""
    return 8.0 * sum(6 * int(d[0]) ** 2 for d in b.dims)
