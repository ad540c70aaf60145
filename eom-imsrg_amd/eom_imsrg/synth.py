"""Seeded synthetic operators of the benchmark configurations.

The reference has no generator for these inputs (its tests draw U(-5,5)
with the compiler's RNG, brute_force_testing.f90:15-120), so the
construction is fixed here once, host-side, and shared byte-for-byte by the
GPU path and the CPU oracle:

RNG     SplitMix64(seed): state += 0x9E3779B97F4A7C15, output mix
        (z^(z>>30))*0xBF58476D1CE4E5B9, (z^(z>>27))*0x94D049BB133111EB,
        z^(z>>31); uniform u = (z >> 11) * 2^-53.
Normals Box-Muller on consecutive uniform pairs (u1, u2):
        r = sqrt(-2 ln(1-u1)); n0 = r cos(2 pi u2), n1 = r sin(2 pi u2).
        A request for k normals consumes 2*ceil(k/2) uniforms.
Order   one-body off-diagonals first (orbit pairs a < b with equal l, 2j,
        2tz, a-major), then every channel q ascending, blocks gam1..gam6;
        square blocks draw their upper triangle (row <= col) column by
        column, rectangular blocks every element column by column.
Pauli   a pair (a,a) at odd J gets zero (its draw is still consumed), as
        construct_random_rank0 (brute_force_testing.f90:99-107).

Hamiltonian (vacuum form, herm +1): t_aa = hw (2n+l+3/2), t_ab = 0.5 N(0,1)
for a != b of one (l,j,tz) block; V stored normalised elements =
N(0,1) exp(-(e_a+e_b+e_c+e_d)/8), square blocks symmetric.  The flowed H is
its normal-ordered form (normal_order, basic_IMSRG.f90:2419-2550).

Omega (herm -1): entries N(0, sigma); square blocks and fhh / fpp
antisymmetric with zero diagonal; fph and rectangular blocks free.
"""
from __future__ import annotations

import math

import numpy as np

from .basis import Basis, SQUARE

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def raw(self, k: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + k + 1, dtype=np.uint64)
        self.count += k
        with np.errstate(over="ignore"):
            z = self.seed + idx * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))

    def uniform(self, k: int) -> np.ndarray:
        return (self.raw(k) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, k: int) -> np.ndarray:
        m = (k + 1) // 2
        u = self.uniform(2 * m)
        r = np.sqrt(-2.0 * np.log1p(-u[0::2]))
        th = 2.0 * math.pi * u[1::2]
        out = np.empty(2 * m)
        out[0::2] = r * np.cos(th)
        out[1::2] = r * np.sin(th)
        return out[:k]


def _same_block_pairs(b: Basis):
    N = b.norb
    pr = []
    for a in range(N):
        for c in range(a + 1, N):
            if b.l[a] == b.l[c] and b.jj[a] == b.jj[c] and b.tz[a] == b.tz[c]:
                pr.append((a, c))
    return pr


def _set_one_body(b: Basis, op: np.ndarray, a: int, c: int, x: float, herm: int) -> None:
    """set f(a,c) = x (and the implied partner f(c,a) = herm*x) in the flat layout"""
    A, P = b.nhole, b.npart
    ia = int(np.searchsorted(b.holes if b.hole[a] else b.parts, a))
    ic = int(np.searchsorted(b.holes if b.hole[c] else b.parts, c))
    ca, cc = b.hole[a], b.hole[c]
    if ca and cc:
        op[b.off_fhh + ia + ic * A] = x
        op[b.off_fhh + ic + ia * A] = herm * x
    elif not ca and not cc:
        op[b.off_fpp + ia + ic * P] = x
        op[b.off_fpp + ic + ia * P] = herm * x
    elif not ca:          # a particle, c hole: fph(a, c)
        op[b.off_fph + ia + ic * P] = x
    else:                 # a hole, c particle: stored fph(c, a) = herm * x
        op[b.off_fph + ic + ia * P] = herm * x


def _fill_two_body(b: Basis, op: np.ndarray, rng: SplitMix64, herm: int, sigma: float, damp: bool) -> None:
    for q in range(b.nblocks):
        J = int(b.jtp[q, 0])
        odd = (J // 2) % 2 == 1
        for g in range(6):
            n1, n2 = b.block_shape(q, g)
            if n1 * n2 == 0:
                continue
            rows = b.pairs[q][(0, 0, 0, 1, 2, 1)[g]]
            cols = b.pairs[q][(0, 1, 2, 1, 2, 2)[g]]
            if SQUARE[g]:
                counts = np.arange(1, n1 + 1)          # column JX holds rows 0..JX
                jx = np.repeat(np.arange(n1), counts)
                ix = np.arange(int(counts.sum())) - np.repeat(np.cumsum(counts) - counts, counts)
            else:
                jx, ix = np.divmod(np.arange(n1 * n2), n1)
            x = rng.normal(len(ix)) * sigma
            a, bb = rows[ix, 0], rows[ix, 1]
            c, d = cols[jx, 0], cols[jx, 1]
            if damp:
                x = x * np.exp(-(b.e[a] + b.e[bb] + b.e[c] + b.e[d]) / 8.0)
            if odd:
                x = np.where((a == bb) | (c == d), 0.0, x)
            blk = np.zeros((n1, n2))
            if SQUARE[g]:
                if herm == -1:
                    x = np.where(ix == jx, 0.0, x)
                blk[ix, jx] = x
                blk[jx, ix] = herm * x
            else:
                blk[ix, jx] = x
            o = int(b.goff[q, g])
            op[o: o + n1 * n2] = blk.reshape(-1, order="F")


def hamiltonian_vacuum(b: Basis, seed: int = 0, hw: float = 20.0, v_scale: float = 1.0,
                       rng: SplitMix64 | None = None) -> np.ndarray:
    """(t, V) of the synthetic Hamiltonian before normal ordering."""
    rng = rng or SplitMix64(seed)
    op = np.zeros(b.oplen)
    for a in range(b.norb):
        _set_one_body(b, op, a, a, hw * (b.e[a] + 1.5), 1)
    pr = _same_block_pairs(b)
    x = 0.5 * rng.normal(len(pr))
    for (a, c), v in zip(pr, x):
        _set_one_body(b, op, a, c, float(v), 1)
    _fill_two_body(b, op, rng, 1, 1.0, True)
    if v_scale != 1.0:
        op[b.off_fpp + b.npart * b.npart:] *= v_scale
    return op


def instance_parameters(k: int):
    """(seed, two-body scale, hw) of instance k of the many-instance batch:
    scale ~ U(0.8, 1.2) and hw in {16, 20, 24} from SplitMix64(1_000_003 + k)"""
    r = SplitMix64(1_000_003 + k)
    u = r.uniform(2)
    return k, 0.8 + 0.4 * float(u[0]), (16.0, 20.0, 24.0)[min(int(3 * u[1]), 2)]


def antihermitian(b: Basis, seed: int = 1, sigma: float = 0.05) -> np.ndarray:
    """random antihermitian operator (herm -1) with N(0, sigma) entries"""
    rng = SplitMix64(seed)
    op = np.zeros(b.oplen)
    pr = []
    for a in range(b.norb):
        for c in range(a + 1, b.norb):
            if b.l[a] == b.l[c] and b.jj[a] == b.jj[c] and b.tz[a] == b.tz[c]:
                pr.append((a, c))
    x = sigma * rng.normal(len(pr))
    for (a, c), v in zip(pr, x):
        _set_one_body(b, op, a, c, float(v), -1)
    _fill_two_body(b, op, rng, -1, sigma, False)
    return op


def hermitian_random(b: Basis, seed: int = 2, sigma: float = 1.0) -> np.ndarray:
    """random hermitian operator (herm +1) with N(0, sigma) entries, E0 = 0"""
    rng = SplitMix64(seed)
    op = np.zeros(b.oplen)
    for a in range(b.norb):
        _set_one_body(b, op, a, a, float(sigma * rng.normal(1)[0]), 1)
    pr = _same_block_pairs(b)
    x = sigma * rng.normal(len(pr))
    for (a, c), v in zip(pr, x):
        _set_one_body(b, op, a, c, float(v), 1)
    _fill_two_body(b, op, rng, 1, sigma, False)
    return op


def me2j_values(orbits: dict, seed: int = 0, scale: float = 2.0) -> np.ndarray:
    """synthetic isospin-coupled two-body matrix elements for an me2j file over
    the pn orbit table `orbits`: N(0,1) * scale * exp(-(e1+e2+e3+e4)/8) per
    isospin component (numpy's default_rng(seed)), with the components a
    pair of identical isospin orbits cannot carry set to zero (two nucleons
    in one j-shell: 2J/2 even needs T = 1, odd needs T = 0) and
    charge-independent T = 1 values (the same for every Mt)"""
    from .interaction import me2j_quadruples
    rng = np.random.default_rng(seed)
    e = 2 * np.asarray(orbits["n"])[::2] + np.asarray(orbits["l"])[::2]
    out = []
    for k1, k2, kk1, kk2, J in me2j_quadruples(orbits):
        w = scale * np.exp(-(e[k1] + e[k2] + e[kk1] + e[kk2]) / 8.0)
        v0, v1 = rng.standard_normal(2) * w
        ident = k1 == k2 or kk1 == kk2
        if ident and (J // 2) % 2 == 0:
            v0 = 0.0
        if ident and (J // 2) % 2 == 1:
            v1 = 0.0
        out.extend((v0, v1, v1, v1))
    return np.array(out)


# tgam g <-> its transposed partner in the J1 = J2 blocks (pppp, phph, hhhh
# are their own; ppph <-> phpp, pphh <-> hhpp, phhh <-> hhph)
_TMIRROR = (0, 7, 6, 3, 4, 8, 2, 1, 5)


def tensor_random(layout, seed: int = 5, herm: int = 1, sigma: float = 1.0) -> np.ndarray:
    """random tensor operator in the flat layout of a tensor.TensorLayout:
    one-body elements N(0, sigma) where rank, parity and tz allow them (fpp
    and fhh with f(q,p) = herm (-1)^((jp-jq)/2) f(p,q), as the commutators
    mirror them); two-body blocks N(0, sigma), Pauli-forbidden pairs zero, and
    in the J1 = J2 blocks <cd J||T||ab J> = herm <ab J||T||cd J> (the
    orientation the storage keeps twice).  Draw order: one-body pairs, then
    blocks q ascending, tgam 1..9 column by column (the mirrored halves are
    overwritten, their draws still consumed)."""
    b = layout.basis
    rng = SplitMix64(seed)
    op = np.zeros(layout.oplen)
    rank, dp = layout.rank, layout.dpar // 2
    l, jj, tz = np.asarray(b.l), np.asarray(b.jj), np.asarray(b.tz)

    def ok1(x, y):
        return (tz[x] == tz[y]) & ((l[x] + l[y] + dp) % 2 == 0) & (np.abs(jj[x] - jj[y]) <= rank) & \
            (rank <= jj[x] + jj[y])

    for lst, off in ((np.asarray(b.holes), b.off_fhh), (np.asarray(b.parts), b.off_fpp)):
        n = len(lst)
        if n == 0:
            continue
        X = sigma * rng.normal(n * n).reshape(n, n, order="F")
        a, c = lst[:, None], lst[None, :]
        ph = herm * (1 - 2 * ((np.abs(jj[a] - jj[c]) // 2) % 2))
        X = np.where(ok1(a, c), X, 0.0)
        up = np.triu(X)
        M = up + (ph * up).T
        np.fill_diagonal(M, np.where(np.diag(ph) > 0, np.diag(up), 0.0))
        op[off: off + n * n] = M.reshape(-1, order="F")
    P, A = len(b.parts), len(b.holes)
    if P * A:
        X = sigma * rng.normal(P * A).reshape(P, A, order="F")
        X = np.where(ok1(np.asarray(b.parts)[:, None], np.asarray(b.holes)[None, :]), X, 0.0)
        op[b.off_fph: b.off_fph + P * A] = X.reshape(-1, order="F")

    def pauli_mask(pairs, J):
        pr = np.asarray(pairs).reshape(-1, 2)
        return ~((pr[:, 0] == pr[:, 1]) & ((J // 2) % 2 == 1))

    blocks = {}
    for q in range(layout.nblocks):
        J1, J2, T, P1, P2, q1, q2 = (int(x) for x in layout.info[q])
        if q2 < 0:
            continue
        for g in range(9):
            n1, n2 = layout.shape(q, g)
            if n1 * n2 == 0:
                continue
            X = sigma * rng.normal(n1 * n2).reshape(n1, n2, order="F")
            rm = pauli_mask(b.pairs[q1][layout.SEA1[g]], J1)
            cm = pauli_mask(b.pairs[q2][layout.SEA2[g]], J2)
            X = X * rm[:, None] * cm[None, :]
            blocks[(q, g)] = X
    look = {}
    for q in range(layout.nblocks):
        J1, J2, T, P1 = (int(x) for x in layout.info[q, :4])
        look[(J1, J2, T, P1)] = q
    for q in range(layout.nblocks):
        J1, J2, T, P1, P2, q1, q2 = (int(x) for x in layout.info[q])
        if q2 < 0 or J1 != J2:
            continue
        qm = look[(J1, J1, T, P2)]
        for g in range(9):
            gm = _TMIRROR[g]
            if (q, g) not in blocks or (qm, gm) not in blocks:
                continue
            if (qm, gm) < (q, g):
                continue
            X = blocks[(q, g)]
            if (qm, gm) == (q, g):
                up = np.triu(X)
                M = up + herm * np.triu(X, 1).T
                if herm < 0:
                    np.fill_diagonal(M, 0.0)
                blocks[(q, g)] = M
            else:
                blocks[(qm, gm)] = herm * X.T
    for (q, g), X in blocks.items():
        o = int(layout.goff[q, g])
        op[o: o + X.size] = X.reshape(-1, order="F")
    return op
