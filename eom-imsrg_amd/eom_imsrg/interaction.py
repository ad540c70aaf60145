"""A physical Hamiltonian on the basis: the harmonic-oscillator one-body part
(calculate_h0_harm_osc, basic_IMSRG.f90:2093-2191) and the two-body matrix
elements of an me2j interaction file (read_me2j_interaction,
interaction_IO.f90:282-557), as the reference's main program sets them up
before normal ordering (main_IMSRG.f90:121-140).

The scalars the reference keeps on the Hamiltonian (hospace, lawson_beta,
com_hw, Aprot + Aneut) are arguments; ``htype`` is the reference's
Hamiltonian type (1: T - Tcm + V, 2: T + Uho + V, 3: T + V).  The vacuum
Hamiltonian is then normal ordered either in the HO basis
(``commutators.normal_order``) or in the Hartree-Fock basis
(``hartree_fock``, calc_HF of HF_mod.f90:14-105).
"""
from __future__ import annotations

import ctypes as C
import gzip

import numpy as np

from . import _lib
from .operators import SqOp, _stream
from .orbits import hk_orbits


def h0_harm_osc(H: SqOp, hw: float, htype: int = 1, mass: int = 16, lawson_beta: float = 0.0,
                com_hw: float | None = None) -> SqOp:
    """fhh, fph, fpp of H <- the HO one-body Hamiltonian of type htype"""
    _lib.check(_lib.lib.imsrg_h0_harm_osc(H.basis.handle, float(hw), int(htype), float(lawson_beta),
                                          float(hw if com_hw is None else com_hw), int(mass), H.ptr, _stream()))
    return H


def file_orbits(file_emax: int, file_lmax: int | None = None) -> dict:
    """orbit table of the .sps file an me2j file of that eMax is written on"""
    return hk_orbits(file_emax, file_lmax)


def read_me2j(H: SqOp, path: str, file_emax: int | None = None, hw: float = 20.0, htype: int = 1, mass: int = 16,
              lawson_beta: float = 0.0, com_hw: float | None = None, orbits: dict | None = None) -> SqOp:
    """two-body blocks of H <- the me2j file's matrix elements (proton-neutron
    form, centre-of-mass and Lawson terms added); ``orbits`` (n, l, jj, tz)
    is the file's orbit table, by default the hk table of ``file_emax``"""
    if orbits is None:
        if file_emax is None:
            raise ValueError("give file_emax or the file's orbit table")
        orbits = file_orbits(file_emax)
    arr = [np.ascontiguousarray(orbits[k], np.int32) for k in ("n", "l", "jj", "tz")]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_int)) for a in arr]
    _lib.check(_lib.lib.imsrg_read_me2j(H.basis.handle, str(path).encode(), len(arr[0]), *ptrs, float(hw), int(htype),
                                        float(lawson_beta), float(hw if com_hw is None else com_hw), int(mass),
                                        H.ptr, _stream()))
    return H


def hartree_fock(H: SqOp, maxit: int = 200):
    """H (vacuum (t, V), HO basis) <- normal ordered in the Hartree-Fock basis;
    returns (e_HF, iterations, HF orbits per (tz, l, j) block, flat
    column-major)"""
    e = C.c_double()
    it = C.c_int()
    n = int(_lib.lib.imsrg_basis_size(H.basis.handle, _lib.SIZE["SPMAT"]))
    D = np.zeros(max(n, 1))
    _lib.check(_lib.lib.imsrg_hartree_fock(H.basis.handle, H.ptr, int(maxit), C.byref(e), C.byref(it),
                                           D.ctypes.data_as(C.POINTER(C.c_double)), _stream()))
    return e.value, it.value, D[:n]


def me2j_quadruples(orbits: dict):
    """(k1, k2, kk1, kk2, 2J) of an me2j file over the isospin orbits of the
    pn table `orbits` (entries 2k proton, 2k+1 neutron), in file order
    (interaction_IO.f90:318-350): k1 >= k2, kk1 <= k1, kk2 <= kk1 (<= k2 when
    kk1 == k1), equal parity, every common 2J; four values each in the order
    (T, Mt) = (0, 0), (1, -1), (1, 0), (1, 1)"""
    l, jj = np.asarray(orbits["l"])[::2], np.asarray(orbits["jj"])[::2]
    nt = len(l)
    for k1 in range(nt):
        for k2 in range(k1 + 1):
            for kk1 in range(k1 + 1):
                top = k2 if kk1 == k1 else kk1
                for kk2 in range(top + 1):
                    if (l[k1] + l[k2]) % 2 != (l[kk1] + l[kk2]) % 2:
                        continue
                    lo = max(abs(jj[k1] - jj[k2]), abs(jj[kk1] - jj[kk2]))
                    hi = min(jj[k1] + jj[k2], jj[kk1] + jj[kk2])
                    for J in range(lo, hi + 1, 2):
                        yield k1, k2, kk1, kk2, J


def write_me2j(path: str, values, compress: bool = True, header: str = "me2j interaction file",
               level: int = 6) -> str:
    """values (in me2j_quadruples order, four per entry) as an me2j file: one
    header line, ten values a line as f12.7 then f13.7, gzip-compressed
    unless told not to (the format read_me2j reads)"""
    v = np.round(np.asarray(values, np.float64), 7)
    lines = [header.replace("\n", " ") + "\n"]
    for i in range(0, len(v), 10):
        chunk = v[i: i + 10]
        lines.append(f"{chunk[0]:12.7f}" + "".join(f"{x:13.7f}" for x in chunk[1:]) + "\n")
    data = "".join(lines).encode()
    if compress:
        with gzip.open(path, "wb", compresslevel=level) as f:
            f.write(data)
    else:
        with open(path, "wb") as f:
            f.write(data)
    return path


def sqop_to_hf(Op: SqOp, coefs: np.ndarray) -> SqOp:
    """sqop_to_HF (HF_mod.f90:181-209): the hermitian scalar operator Op
    (vacuum form, HO basis) <- rotated to the HF orbits `coefs` (as
    hartree_fock returns them) and normal ordered"""
    D = np.ascontiguousarray(coefs, np.float64).reshape(-1)
    n = int(_lib.lib.imsrg_basis_size(Op.basis.handle, _lib.SIZE["SPMAT"]))
    if D.size != n:
        raise ValueError(f"coefs holds {D.size} doubles; the basis's HF orbit blocks need {n}")
    _lib.check(_lib.lib.imsrg_sqop_to_hf(Op.basis.handle, D.ctypes.data_as(C.POINTER(C.c_double)), Op.ptr,
                                         int(Op.herm), _stream()))
    return Op
