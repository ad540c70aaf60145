"""J-scheme helpers of the CPU tests: dense one-body matrices and element-wise
construction of operators in the flat layout, through the oracle's f_elem /
v_elem (test infrastructure; no product code path)."""
import numpy as np

ROWS = (0, 0, 0, 1, 2, 1)   # pair class of the rows of gam1..gam6 (pp, ph, hh)
COLS = (0, 1, 2, 1, 2, 2)


def one_body(o, b, op, herm):
    N = b.norb
    return np.array([[o.f_elem(op, herm, a, c) for c in range(N)] for a in range(N)])


def set_one_body(b, op, F):
    """write the N x N one-body matrix F into fhh / fph / fpp"""
    A, P = b.nhole, b.npart
    H, Pp = b.holes, b.parts
    op[b.off_fhh: b.off_fhh + A * A] = F[np.ix_(H, H)].reshape(-1, order="F")
    op[b.off_fph: b.off_fph + P * A] = F[np.ix_(Pp, H)].reshape(-1, order="F")
    op[b.off_fpp: b.off_fpp + P * P] = F[np.ix_(Pp, Pp)].reshape(-1, order="F")


def each_element(b):
    """(q, g, ix, jx, (a, b), (c, d), 2J) for every stored two-body element"""
    for q in range(b.nblocks):
        J = int(b.jtp[q, 0])
        for g in range(6):
            n1, n2 = b.block_shape(q, g)
            if n1 * n2 == 0:
                continue
            rows, cols = b.pairs[q][ROWS[g]], b.pairs[q][COLS[g]]
            for jx in range(n2):
                for ix in range(n1):
                    yield q, g, ix, jx, tuple(rows[ix]), tuple(cols[jx]), J


def set_two_body(b, op, value):
    """stored (normalised) element = value(a,b,c,d,2J) / sqrt(1+d_ab) / sqrt(1+d_cd)"""
    for q, g, ix, jx, (a, bb), (c, d), J in each_element(b):
        n1, _ = b.block_shape(q, g)
        v = value(a, bb, c, d, J)
        if a == bb:
            v /= np.sqrt(2.0)
        if c == d:
            v /= np.sqrt(2.0)
        op[int(b.goff[q, g]) + ix + jx * n1] = v


def two_body_part(b):
    return slice(b.off_fpp + b.npart * b.npart, b.oplen)


def ljt_blocks(b):
    """orbit index lists of equal (l, 2j, 2tz)"""
    out = {}
    for a in range(b.norb):
        out.setdefault((int(b.l[a]), int(b.jj[a]), int(b.tz[a])), []).append(a)
    return list(out.values())


def random_tensor(ot, b, herm, seed, sigma=1.0):
    """random tensor operator in the flat tensor layout of `ot` (an
    OracleTensor over basis b): selection rules of rank / parity / tz on the
    one-body part, Pauli-forbidden elements zero, and the hermiticity
    <cd J||T||ab J> = herm <ab J||T||cd J>, f(q,p) = herm (-1)^((jp-jq)/2)
    f(p,q) that the reference's storage (both J1 = J2 orientations, full
    fpp / fhh) assumes"""
    rng = np.random.default_rng(seed)
    op = ot.zeros()
    A, P = b.nhole, b.npart
    H, Pp = list(b.holes), list(b.parts)

    def ok1(a, c):
        return (b.tz[a] == b.tz[c] and (b.l[a] + b.l[c] + ot.dpar // 2) % 2 == 0
                and abs(b.jj[a] - b.jj[c]) <= ot.rank <= b.jj[a] + b.jj[c])

    for lst, off, n in ((H, b.off_fhh, A), (Pp, b.off_fpp, P)):
        for i in range(n):
            for k in range(i, n):
                a, c = lst[i], lst[k]
                if not ok1(a, c):
                    continue
                v = sigma * rng.normal()
                ph = herm * (1 - 2 * ((abs(int(b.jj[a]) - int(b.jj[c])) // 2) % 2))
                if i == k and ph < 0:
                    v = 0.0
                op[off + i + k * n] = v
                op[off + k + i * n] = ph * v
    for i, a in enumerate(Pp):
        for k, c in enumerate(H):
            if ok1(a, c):
                op[b.off_fph + i + k * P] = sigma * rng.normal()
    done = {}
    for q, g, ix, jx, p1, p2, J1, J2, o in ot.elements():
        if (p1[0] == p1[1] and (J1 // 2) % 2) or (p2[0] == p2[1] and (J2 // 2) % 2):
            continue
        if J1 == J2 and (p2, p1, J1) in done:
            op[o] = herm * op[done[(p2, p1, J1)]]
            continue
        if J1 == J2 and p1 == p2 and herm < 0:
            continue
        op[o] = sigma * rng.normal()
        done[(p1, p2, J1)] = o
    return op
