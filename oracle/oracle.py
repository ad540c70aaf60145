"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle
(oracle/imsrg_oracle.cpp, built into oracle/_build/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  Operators are numpy float64 arrays in the same flat
layout as the product (E0, fhh, fph, fpp, then gam1..6 per channel).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp, d, i, dp, ip, lp = C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), \
            C.POINTER(C.c_int64)
        sig = {
            "orc_set_blas": (i, [C.c_char_p]), "orc_unset_blas": (None, []), "orc_set_threads": (i, [i]),
            "orc_basis_new": (vp, [i, ip, ip, ip, ip, ip]), "orc_basis_free": (None, [vp]),
            "orc_basis_sizes": (None, [vp, lp]), "orc_basis_channels": (None, [vp, ip, lp, ip]),
            "orc_basis_pairs": (i, [vp, i, i, ip]), "orc_basis_cc": (None, [vp, ip, ip, lp, lp, ip]),
            "orc_cc_row": (i, [vp, i, i, i]), "orc_cc_col": (i, [vp, i, i, i]),
            "orc_sixj": (d, [i] * 6), "orc_ninej": (d, [i] * 9), "orc_sixj_table": (d, [vp] + [i] * 6),
            "orc_v_elem": (d, [vp, dp, i, i, i, i, i, i]), "orc_f_elem": (d, [vp, dp, i, i, i]),
            "orc_comm110": (d, [vp, dp, i, dp, i]), "orc_comm220": (d, [vp, dp, i, dp, i]),
            "orc_comm111": (None, [vp, dp, i, dp, i, dp, i]), "orc_comm121": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_comm122": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_comm222pphh": (None, [vp, dp, i, dp, i, dp, i, dp, dp]),
            "orc_comm221": (None, [vp, dp, i, dp, i, dp, i, dp, dp]),
            "orc_cross_coupled": (None, [vp, dp, i, dp]),
            "orc_comm222ph": (None, [vp, dp, i, dp, i, dp, i, dp]),
            "orc_frob_norm": (d, [vp, dp]), "orc_add": (None, [vp, dp, d, dp, d, dp]),
            "orc_white": (None, [vp, dp, dp]), "orc_mbpt2": (d, [vp, dp]), "orc_normal_order": (None, [vp, dp]),
            "orc_monopole": (d, [vp, dp, i, i]),
            "orc_commutator": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_bch": (i, [vp, dp, dp, dp, d, i, dp]), "orc_magnus": (i, [vp, dp, dp, d]),
            "orc_magnus_decouple": (i, [vp, dp, dp, d, d, d, i, d, i, dp]),
            "orc_magnus_decouple_gen": (i, [vp, dp, dp, d, d, d, i, d, i, dp, i]),
            "orc_gs_atan": (None, [vp, dp, dp]), "orc_gs_atanxx": (None, [vp, dp, dp, dp]),
            "orc_gs_w2": (None, [vp, dp, dp]),
            "orc_commutator_vjp": (None, [vp, dp, i, dp, i, dp, i, dp, dp]),
            "orc_bch_vjp": (i, [vp, dp, dp, d, i, dp, dp, dp]),
            "orc_magnus_vjp": (i, [vp, dp, dp, d, dp, dp, dp]),
            "orc_white_vjp": (None, [vp, dp, dp, dp]),
            "orc_normal_order_vjp": (None, [vp, dp, dp]),
            "orc_flow_grad": (i, [vp, dp, d, d, d, d, i, i, dp, dp]),
            "orc_flow_grad_gen": (i, [vp, dp, d, d, d, d, i, i, dp, dp, i]),
            "orc_gen_vjp": (None, [vp, dp, dp, dp, dp, dp, i]), "orc_gs_w2_vjp": (None, [vp, dp, dp, dp]),
            "orc_bf_0body": (d, [vp, dp, i, dp, i]), "orc_bf_1body": (d, [vp, dp, i, dp, i, i, i]),
            "orc_bf_2body": (d, [vp, dp, i, dp, i, i, i, i, i, i, dp]),
            "orc_h0_harm_osc": (None, [vp, d, i, d, d, i, dp]),
            "orc_me2j_count": (C.c_int64, [i, ip, ip]),
            "orc_me2j_convert": (i, [vp, dp, C.c_int64, i, ip, ip, ip, d, i, d, d, i, dp]),
            "orc_r1_r2": (d, [vp, i, i, i, i, i]), "orc_p1_p2": (d, [vp, i, i, i, i, i]),
            "orc_sp_layout": (i, [vp, ip]), "orc_hf_kin": (None, [vp, dp, dp]),
            "orc_hf_density": (None, [vp, dp, dp]), "orc_hf_gamma": (None, [vp, dp, dp, dp]),
            "orc_hf_finish": (None, [vp, dp, dp, dp, dp, dp]), "orc_hf_transform_2b": (None, [vp, dp, dp]),
            "orc_sqop_to_hf": (None, [vp, dp, dp]),
            "orc_tensor_new": (vp, [vp, i, i]), "orc_tensor_free": (None, [vp]),
            "orc_tensor_sizes": (None, [vp, lp]), "orc_tensor_blocks": (None, [vp, ip, ip, lp]),
            "orc_tensor_block_index_ref": (i, [i] * 5), "orc_xxxsixj": (d, [i] * 6),
            "orc_ninej_ref": (d, [i] * 9),
            "orc_tensor_elem": (d, [vp, dp, i, i, i, i, i, i, i]), "orc_f_tensor_elem": (d, [vp, dp, i, i, i]),
            "orc_ts111": (None, [vp, dp, i, dp, i, dp, i]), "orc_ts121": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_ts211": (None, [vp, dp, i, dp, i, dp, i]), "orc_ts122": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_ts212": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_ts222pphh": (None, [vp, dp, i, dp, i, dp, i, dp, dp]),
            "orc_ts221": (None, [vp, dp, dp, i, dp, i]), "orc_ts222ph": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_tensor_pandya": (None, [vp, dp, i, dp]),
            "orc_ts_commutator": (None, [vp, dp, i, dp, i, dp, i]),
            "orc_bch_tensor": (i, [vp, dp, i, dp, i, d, dp]), "orc_tensor_frob_norm": (d, [vp, dp]),
            "orc_copy_rank0_to_tensor": (None, [vp, dp, i, dp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def use_host_blas() -> str | None:
    """route the oracle's dgemm to numpy's bundled OpenBLAS when present"""
    import glob
    import numpy
    libs = glob.glob(os.path.join(os.path.dirname(numpy.__file__) + ".libs", "libscipy_openblas64_*.so"))
    for p in libs:
        if lib().orc_set_blas(p.encode()) == 0:
            return p
    return None


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleBasis:
    def __init__(self, n, l, jj, tz, hole):
        ip = lambda a: np.ascontiguousarray(a, np.int32).ctypes.data_as(C.POINTER(C.c_int))
        self._keep = [np.ascontiguousarray(x, np.int32) for x in (n, l, jj, tz, hole)]
        self.h = lib().orc_basis_new(len(n), *[x.ctypes.data_as(C.POINTER(C.c_int)) for x in self._keep])
        s = np.zeros(9, np.int64)
        lib().orc_basis_sizes(self.h, s.ctypes.data_as(C.POINTER(C.c_int64)))
        (self.N, self.A, self.P, self.JTM, self.nblocks, self.oplen, self.ncc, self.cclen,
         self.wcclen) = [int(x) for x in s]
        nb = self.nblocks
        self.dims = np.zeros(3 * nb, np.int32)
        self.goff = np.zeros(6 * nb, np.int64)
        self.jtp = np.zeros(3 * nb, np.int32)
        lib().orc_basis_channels(self.h, self.dims.ctypes.data_as(C.POINTER(C.c_int)),
                                 self.goff.ctypes.data_as(C.POINTER(C.c_int64)),
                                 self.jtp.ctypes.data_as(C.POINTER(C.c_int)))
        self.dims = self.dims.reshape(nb, 3)
        self.goff = self.goff.reshape(nb, 6)
        self.jtp = self.jtp.reshape(nb, 3)

    def pairs(self, q, c):
        buf = np.zeros(2 * max(int(self.dims[q, c]), 1), np.int32)
        k = lib().orc_basis_pairs(self.h, q, c, buf.ctypes.data_as(C.POINTER(C.c_int)))
        return buf[: 2 * k].reshape(k, 2)

    def cc_layout(self):
        m = self.ncc
        rl, nc = np.zeros(m, np.int32), np.zeros(m, np.int32)
        co, wo, jtp = np.zeros(m, np.int64), np.zeros(m, np.int64), np.zeros(3 * m, np.int32)
        lib().orc_basis_cc(self.h, rl.ctypes.data_as(C.POINTER(C.c_int)), nc.ctypes.data_as(C.POINTER(C.c_int)),
                           co.ctypes.data_as(C.POINTER(C.c_int64)), wo.ctypes.data_as(C.POINTER(C.c_int64)),
                           jtp.ctypes.data_as(C.POINTER(C.c_int)))
        return rl, nc, co, wo, jtp.reshape(m, 3)

    # --- vectorize / repackage, basic_IMSRG.f90:3929-4053, on the flat layout
    # [E0 | fhh A x A | fph P x A | fpp P x P | channel blocks]
    def _one_body_views(self, op):
        A, P = self.A, self.P
        o1 = 1 + A * A
        o2 = o1 + P * A
        return op[1:o1], op[o2: o2 + P * P], op[o1:o2]   # fhh, fpp, fph (column-major)

    def _blocks(self):
        for q in range(self.nblocks):
            npp, nph, nhh = (int(x) for x in self.dims[q])
            shapes = [(npp, npp), (npp, nph), (npp, nhh), (nph, nph), (nhh, nhh), (nph, nhh)]
            for g, (nr, nc) in enumerate(shapes):
                yield int(self.goff[q, g]), nr, nc, g in (0, 3, 4)

    def vectorize(self, op):
        out = [op[:1], *self._one_body_views(op)]
        for o, nr, nc, square in self._blocks():
            X = op[o: o + nr * nc].reshape(nc, nr)        # X[i] = column i
            if square:
                out.extend(X[i, i:] for i in range(nc))   # rows j = i .. nr-1
            else:
                out.append(X.reshape(-1))
        return np.concatenate(out)

    def repackage(self, vec, herm):
        op = np.zeros(self.oplen)
        fhh, fpp, fph = self._one_body_views(op)
        A, P = self.A, self.P
        op[0] = vec[0]
        k = 1
        for part, n in ((fhh, A * A), (fpp, P * P), (fph, P * A)):
            part[:] = vec[k: k + n]
            k += n
        for o, nr, nc, square in self._blocks():
            X = op[o: o + nr * nc].reshape(nc, nr)        # view: X[i, j] = element (j, i)
            for i in range(nc):
                j0 = i if square else 0
                for j in range(j0, nr):
                    X[i, j] = vec[k]
                    if square:
                        X[j, i] = vec[k] * herm
                    k += 1
        assert k == len(vec)
        return op

    # --- reference routines on host arrays
    def v_elem(self, op, herm, a, b, c, d, J):
        return lib().orc_v_elem(self.h, _dp(op), herm, a, b, c, d, J)

    def f_elem(self, op, herm, a, b):
        return lib().orc_f_elem(self.h, _dp(op), herm, a, b)

    def comm110(self, L, Lh, R, Rh):
        return lib().orc_comm110(self.h, _dp(L), Lh, _dp(R), Rh)

    def comm220(self, L, Lh, R, Rh):
        return lib().orc_comm220(self.h, _dp(L), Lh, _dp(R), Rh)

    def comm111(self, L, Lh, R, Rh, RES, RESh):
        lib().orc_comm111(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh)

    def comm121(self, L, Lh, R, Rh, RES, RESh):
        lib().orc_comm121(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh)

    def comm122(self, L, Lh, R, Rh, RES, RESh):
        lib().orc_comm122(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh)

    def comm222pphh(self, L, Lh, R, Rh, RES, RESh, w1, w2):
        lib().orc_comm222pphh(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh, _dp(w1), _dp(w2))

    def comm221(self, L, Lh, R, Rh, RES, RESh, w1, w2):
        lib().orc_comm221(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh, _dp(w1), _dp(w2))

    def cross_coupled(self, op, herm):
        cc = np.zeros(max(self.cclen, 1))
        lib().orc_cross_coupled(self.h, _dp(op), herm, _dp(cc))
        return cc

    def comm222ph(self, LCC, Lh, RCC, Rh, RES, RESh, WCC=None):
        W = WCC if WCC is not None else np.zeros(max(self.wcclen, 1))
        lib().orc_comm222ph(self.h, _dp(LCC), Lh, _dp(RCC), Rh, _dp(RES), RESh, _dp(W))
        return W

    def frob_norm(self, op):
        return lib().orc_frob_norm(self.h, _dp(op))

    def add(self, A, ax, B, bx):
        C_ = np.zeros(self.oplen)
        lib().orc_add(self.h, _dp(A), ax, _dp(B), bx, _dp(C_))
        return C_

    def white(self, H):
        eta = np.zeros(self.oplen)
        lib().orc_white(self.h, _dp(H), _dp(eta))
        return eta

    def gs_atan(self, H):
        eta = np.zeros(self.oplen)
        lib().orc_gs_atan(self.h, _dp(H), _dp(eta))
        return eta

    def gs_atanxx(self, Hold, H):
        eta = np.zeros(self.oplen)
        lib().orc_gs_atanxx(self.h, _dp(Hold), _dp(H), _dp(eta))
        return eta

    def gs_w2(self, H):
        eta = np.zeros(self.oplen)
        lib().orc_gs_w2(self.h, _dp(H), _dp(eta))
        return eta

    def mbpt2(self, H):
        return lib().orc_mbpt2(self.h, _dp(H))

    def normal_order(self, H):
        out = H.copy()
        lib().orc_normal_order(self.h, _dp(out))
        return out

    def monopole(self, H, a, c):
        return lib().orc_monopole(self.h, _dp(H), a, c)

    def commutator(self, L, Lh, R, Rh, RESh):
        RES = np.zeros(self.oplen)
        lib().orc_commutator(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh)
        return RES

    def bch(self, G, H, conv, maxit=40):
        HS = np.zeros(self.oplen)
        adv = np.zeros(maxit)
        it = lib().orc_bch(self.h, _dp(HS), _dp(G), _dp(H), conv, maxit, _dp(adv))
        return HS, it, adv

    def magnus(self, DG, G, conv):
        out = DG.copy()
        n = lib().orc_magnus(self.h, _dp(out), _dp(G), conv)
        return out, n

    def magnus_decouple(self, H, ds, bch_conv, mag_conv, crit_mode, crit_tol, maxsteps=200, generator=0):
        """generator 0 white, 1 atan, 2 w2"""
        HS = H.copy()
        G = np.zeros(self.oplen)
        log = np.zeros(4 * maxsteps)
        st = lib().orc_magnus_decouple_gen(self.h, _dp(HS), _dp(G), ds, bch_conv, mag_conv, crit_mode, crit_tol,
                                           maxsteps, _dp(log), int(generator))
        return HS, G, log[: 4 * max(st, 0)].reshape(-1, 4), st

    # ---- reverse-mode derivatives (flat-array gradients)
    def commutator_vjp(self, L, Lh, R, Rh, RESb, RESh):
        Lb, Rb = np.zeros(self.oplen), np.zeros(self.oplen)
        lib().orc_commutator_vjp(self.h, _dp(L), Lh, _dp(R), Rh, _dp(np.ascontiguousarray(RESb)), RESh, _dp(Lb),
                                 _dp(Rb))
        return Lb, Rb

    def bch_vjp(self, G, H, conv, HSb, maxit=40):
        Gb, Hb = np.zeros(self.oplen), np.zeros(self.oplen)
        it = lib().orc_bch_vjp(self.h, _dp(G), _dp(H), conv, maxit, _dp(np.ascontiguousarray(HSb)), _dp(Gb), _dp(Hb))
        return Gb, Hb, it

    def magnus_vjp(self, ETA, G, conv, DGb):
        Eb, Gb = np.zeros(self.oplen), np.zeros(self.oplen)
        k = lib().orc_magnus_vjp(self.h, _dp(ETA), _dp(G), conv, _dp(np.ascontiguousarray(DGb)), _dp(Eb), _dp(Gb))
        return Eb, Gb, k

    def white_vjp(self, H, ETAb):
        Hb = np.zeros(self.oplen)
        lib().orc_white_vjp(self.h, _dp(H), _dp(np.ascontiguousarray(ETAb)), _dp(Hb))
        return Hb

    def normal_order_vjp(self, Hb):
        Tb = np.zeros(self.oplen)
        lib().orc_normal_order_vjp(self.h, _dp(np.ascontiguousarray(Hb)), _dp(Tb))
        return Tb

    def gen_vjp(self, Hd, Hn, ETAb, mode):
        """(Hdb, Hnb) of gen_core (mode 0 white, 1 atan / atanxx)"""
        Hdb, Hnb = np.zeros(self.oplen), np.zeros(self.oplen)
        lib().orc_gen_vjp(self.h, _dp(Hd), _dp(Hn), _dp(np.ascontiguousarray(ETAb)), _dp(Hdb), _dp(Hnb), int(mode))
        return Hdb, Hnb

    def gs_w2_vjp(self, H, ETAb):
        Hb = np.zeros(self.oplen)
        lib().orc_gs_w2_vjp(self.h, _dp(H), _dp(np.ascontiguousarray(ETAb)), _dp(Hb))
        return Hb

    def flow_grad(self, T, ds, bch_conv, mag_conv, crit_tol, maxsteps=200, ckpt=4, generator=0):
        """(E0, dE0/dT, steps) for magnus_decouple started from normal_order(T)
        (generator 0 white, 1 atan, 2 w2)"""
        Tb = np.zeros(self.oplen)
        e0 = np.zeros(1)
        st = lib().orc_flow_grad_gen(self.h, _dp(T), ds, bch_conv, mag_conv, crit_tol, maxsteps, ckpt, _dp(Tb),
                                     _dp(e0), int(generator))
        return float(e0[0]), Tb, st

    def bf_0body(self, A, Ah, B, Bh):
        return lib().orc_bf_0body(self.h, _dp(A), Ah, _dp(B), Bh)

    def bf_1body(self, A, Ah, B, Bh, a, b):
        return lib().orc_bf_1body(self.h, _dp(A), Ah, _dp(B), Bh, a, b)

    def bf_2body(self, A, Ah, B, Bh, a, b, c, d, J):
        parts = np.zeros(3)
        v = lib().orc_bf_2body(self.h, _dp(A), Ah, _dp(B), Bh, a, b, c, d, J, _dp(parts))
        return v, parts

    # --- Hamiltonian set-up (oracle/imsrg_interaction.inc)
    def h0_harm_osc(self, hw, htype, mass, beta=0.0, cmhw=None, op=None):
        """calculate_h0_harm_osc, basic_IMSRG.f90:2093-2191"""
        op = np.zeros(self.oplen) if op is None else op
        lib().orc_h0_harm_osc(self.h, hw, htype, beta, hw if cmhw is None else cmhw, mass, _dp(op))
        return op

    def r1_r2(self, a, b, c, d, J):
        return lib().orc_r1_r2(self.h, a, b, c, d, J)

    def p1_p2(self, a, b, c, d, J):
        return lib().orc_p1_p2(self.h, a, b, c, d, J)

    def read_me2j(self, path, orbits, hw, htype, mass, beta=0.0, cmhw=None, op=None):
        """read_me2j_interaction, interaction_IO.f90:282-557: the file parsed
        here (gzip or text, fixed-width f12.7 / f13.7 fields), converted by
        the oracle; `orbits` = the file's (n, l, jj, tz) table"""
        fl = np.ascontiguousarray(orbits["l"], np.int32)
        fj = np.ascontiguousarray(orbits["jj"], np.int32)
        nfile = len(fl)
        con = np.full(nfile, -1, np.int32)
        mine = {(int(n), int(l), int(j), int(t)): k for k, (n, l, j, t) in
                enumerate(zip(self._keep[0], self._keep[1], self._keep[2], self._keep[3]))}
        for x in range(nfile):
            con[x] = mine.get((int(orbits["n"][x]), int(fl[x]), int(fj[x]), int(orbits["tz"][x])), -1)
        ipt = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        count = lib().orc_me2j_count(nfile, ipt(fl), ipt(fj))
        me = read_me2j_values(path, count)
        op = np.zeros(self.oplen) if op is None else op
        rc = lib().orc_me2j_convert(self.h, _dp(me), count, nfile, ipt(fl), ipt(fj), ipt(con), hw, htype, beta,
                                    hw if cmhw is None else cmhw, mass, _dp(op))
        if rc != 0:
            raise ValueError("value count does not match the file orbits")
        return op

    # --- Hartree-Fock (oracle/imsrg_hf.inc; the SCF loop of calc_HF,
    # HF_mod.f90:14-105, with numpy's LAPACK eigh as dsyev)
    def sp_blocks(self):
        """sizes of the (tz, l, j) blocks in allocate_sp_mat order and their
        offsets in the flat block storage"""
        nb = lib().orc_sp_layout(self.h, None)
        sz = np.zeros(nb, np.int32)
        lib().orc_sp_layout(self.h, sz.ctypes.data_as(C.POINTER(C.c_int)))
        off = np.concatenate([[0], np.cumsum(sz.astype(np.int64) ** 2)])
        return sz, off

    def hf_transform_2b(self, H, D):
        out = H.copy()
        lib().orc_hf_transform_2b(self.h, _dp(out), _dp(np.ascontiguousarray(D, np.float64)))
        return out

    def sqop_to_hf(self, O, D):
        """sqop_to_HF, HF_mod.f90:181-209 (hermitian scalar operator)"""
        out = O.copy()
        lib().orc_sqop_to_hf(self.h, _dp(out), _dp(np.ascontiguousarray(D, np.float64)))
        return out

    def hartree_fock(self, H):
        """calc_HF (HF_mod.f90:14-105) without three-body forces: returns
        (H in the normal-ordered HF basis, D, iterations); H = the vacuum
        (t, V) in the HO basis"""
        sz, off = self.sp_blocks()
        tot = int(off[-1])
        T = np.zeros(tot)
        lib().orc_hf_kin(self.h, _dp(H), _dp(T))
        D = np.zeros(tot)
        ev = [np.zeros(n) for n in sz]
        for q, n in enumerate(sz):
            D[off[q]: off[q + 1]] = np.eye(n).reshape(-1)
        crit, it = 10.0, 0
        rho, Vg = np.zeros(tot), np.zeros(tot)
        while crit > float(np.float32(1e-6)):
            lib().orc_hf_density(self.h, _dp(D), _dp(rho))
            lib().orc_hf_gamma(self.h, _dp(H), _dp(rho), _dp(Vg))
            F = T + Vg + 0.0
            crit = 0.0
            for q, n in enumerate(sz):
                if n == 0:
                    crit += 0.0
                    continue
                w, v = np.linalg.eigh(F[off[q]: off[q + 1]].reshape(n, n).T)
                D[off[q]: off[q + 1]] = v.T.reshape(-1)          # column-major
                crit += np.sqrt(np.sum((ev[q] - w) ** 2))
                ev[q] = w
            it += 1
        for q, n in enumerate(sz):
            M = D[off[q]: off[q + 1]].reshape(n, n).T.copy()
            for i in range(n):
                if M[i, i] < 0.0:
                    M[:, i] = M[:, i] * (-1)
            D[off[q]: off[q + 1]] = M.T.reshape(-1)
        F = T + Vg + 0.0
        out = H.copy()
        lib().orc_hf_finish(self.h, _dp(out), _dp(D), _dp(F), _dp(T), _dp(Vg))
        return out, D, it

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_basis_free(self.h)
            self.h = None


class OracleTensor:
    """tensor operator layout of rank `rank` (2 lambda) and parity change
    `dpar` (0 or 2) over an OracleBasis (allocate_tensor,
    basic_IMSRG.f90:688-951), with the scalar-tensor commutator terms"""

    def __init__(self, ob: OracleBasis, rank: int, dpar: int = 0):
        self.ob = ob
        self.rank, self.dpar = int(rank), int(dpar)
        self.h = lib().orc_tensor_new(ob.h, self.rank, self.dpar)
        s = np.zeros(5, np.int64)
        lib().orc_tensor_sizes(self.h, s.ctypes.data_as(C.POINTER(C.c_int64)))
        self.oplen, self.nblocks, self.ncc, self.cclen, self.wlen = (int(x) for x in s)
        nb = self.nblocks
        info, dims, goff = np.zeros(7 * nb, np.int32), np.zeros(6 * nb, np.int32), np.zeros(9 * nb, np.int64)
        lib().orc_tensor_blocks(self.h, info.ctypes.data_as(C.POINTER(C.c_int)),
                                dims.ctypes.data_as(C.POINTER(C.c_int)), goff.ctypes.data_as(C.POINTER(C.c_int64)))
        self.info, self.dims, self.goff = info.reshape(nb, 7), dims.reshape(nb, 6), goff.reshape(nb, 9)

    SEA1 = (0, 0, 0, 1, 2, 1, 2, 1, 2)
    SEA2 = (0, 1, 2, 1, 2, 2, 0, 0, 1)

    def shape(self, q, g):
        return int(self.dims[q, self.SEA1[g]]), int(self.dims[q, 3 + self.SEA2[g]])

    def elements(self):
        """(q, g, ix, jx, (a, b), (c, d), 2J1, 2J2, flat offset) of every stored element"""
        for q in range(self.nblocks):
            J1, J2, _, _, _, q1, q2 = (int(x) for x in self.info[q])
            if q2 < 0:
                continue
            for g in range(9):
                n1, n2 = self.shape(q, g)
                if n1 * n2 == 0:
                    continue
                rows, cols = self.ob.pairs(q1, self.SEA1[g]), self.ob.pairs(q2, self.SEA2[g])
                for jx in range(n2):
                    for ix in range(n1):
                        yield q, g, ix, jx, tuple(rows[ix]), tuple(cols[jx]), J1, J2, \
                            int(self.goff[q, g]) + ix + jx * n1

    def zeros(self):
        return np.zeros(self.oplen)

    def tensor_elem(self, op, herm, a, b, c, d, J1, J2):
        return lib().orc_tensor_elem(self.h, _dp(op), herm, a, b, c, d, J1, J2)

    def f_tensor_elem(self, op, herm, a, b):
        return lib().orc_f_tensor_elem(self.h, _dp(op), herm, a, b)

    def _term(self, name, L, Lh, R, Rh, RES, RESh):
        getattr(lib(), name)(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh)
        return RES

    def ts111(self, *a): return self._term("orc_ts111", *a)
    def ts121(self, *a): return self._term("orc_ts121", *a)
    def ts211(self, *a): return self._term("orc_ts211", *a)
    def ts122(self, *a): return self._term("orc_ts122", *a)
    def ts212(self, *a): return self._term("orc_ts212", *a)
    def ts222ph(self, *a): return self._term("orc_ts222ph", *a)

    def ts222pphh(self, L, Lh, R, Rh, RES, RESh, w1=None, w2=None):
        w1 = self.zeros() if w1 is None else w1
        w2 = self.zeros() if w2 is None else w2
        lib().orc_ts222pphh(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh, _dp(w1), _dp(w2))
        return RES, w1, w2

    def ts221(self, w1, w2, pm, RES, RESh):
        lib().orc_ts221(self.h, _dp(w1), _dp(w2), pm, _dp(RES), RESh)
        return RES

    def pandya(self, R, Rh):
        CC = np.zeros(self.cclen)
        lib().orc_tensor_pandya(self.h, _dp(R), Rh, _dp(CC))
        return CC

    def commutator(self, L, Lh, R, Rh, RESh):
        RES = self.zeros()
        lib().orc_ts_commutator(self.h, _dp(L), Lh, _dp(R), Rh, _dp(RES), RESh)
        return RES

    def bch(self, G, Gh, HS, HSh, conv=9.999999974752427e-07):
        out = np.ascontiguousarray(HS, np.float64).copy()
        adv = np.zeros(30)
        it = lib().orc_bch_tensor(self.h, _dp(G), Gh, _dp(out), HSh, conv, _dp(adv))
        return out, it, adv

    def frob_norm(self, op):
        return lib().orc_tensor_frob_norm(self.h, _dp(op))

    def from_rank0(self, A, Ah):
        B = self.zeros()
        lib().orc_copy_rank0_to_tensor(self.h, _dp(A), Ah, _dp(B))
        return B

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tensor_free(self.h)
            self.h = None


def block_index_ref(J1, J2, rank, T, P):
    return lib().orc_tensor_block_index_ref(J1, J2, rank, T, P)


def xxxsixj(*a):
    return lib().orc_xxxsixj(*a)


def ninej_ref(*a):
    return lib().orc_ninej_ref(*a)


def sixj(*a):
    return lib().orc_sixj(*a)


def ninej(*a):
    return lib().orc_ninej(*a)


def read_me2j_values(path, count):
    """the values of an me2j file (interaction_IO.f90:375-392): one header
    line, then lines of fixed-width fields f12.7, 9 x f13.7"""
    import gzip
    opener = gzip.open if open(path, "rb").read(2) == b"\x1f\x8b" else open
    out = np.zeros(count)
    with opener(path, "rt") as f:
        f.readline()
        k = 0
        while k < count:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: ends after {k} of {count} values")
            line = line.rstrip("\n")
            for x in range(min(10, count - k)):
                at, w = (0, 12) if x == 0 else (12 + 13 * (x - 1), 13)
                field = line[at: at + w].strip()
                out[k + x] = float(field) if field else 0.0
            k += min(10, count - k)
    return out
