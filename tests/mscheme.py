"""m-scheme evaluation of the IMSRG(2) commutator, the independent check of
the J-scheme scalar and scalar-tensor commutator formulas (test
infrastructure; no product code path).

Operators are expanded from their reduced J-scheme elements into dense
antisymmetrised m-scheme arrays with Clebsch-Gordan coefficients (Edmonds'
Wigner-Eckart convention: <j1 m1|T^lam_mu|j2 m2> = <j2 m2 lam mu|j1 m1>
<j1||T||j2> / sqrt(2 j1 + 1); a scalar's reduced element is its element
times sqrt(2J + 1)), commuted with the normal-ordered two-body commutator
formulas (one-body and two-body parts, occupations n = 0/1 of the reference
state), and projected back onto reduced elements.  All angular momenta are
doubled integers, as in the J-scheme code.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np


@lru_cache(maxsize=None)
def cg(j1, m1, j2, m2, J, M):
    """<j1 m1 j2 m2 | J M> by Racah's closed formula (doubled arguments)"""
    if m1 + m2 != M or abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0.0
    if J > j1 + j2 or J < abs(j1 - j2) or (j1 + j2 + J) % 2:
        return 0.0
    if (j1 + m1) % 2 or (j2 + m2) % 2 or (J + M) % 2:
        return 0.0
    f = math.factorial
    a, b, c = (j1 + j2 - J) // 2, (j1 - j2 + J) // 2, (-j1 + j2 + J) // 2
    pre = (J + 1) * f(a) * f(b) * f(c) / f((j1 + j2 + J) // 2 + 1)
    pre *= f((j1 + m1) // 2) * f((j1 - m1) // 2) * f((j2 + m2) // 2) * f((j2 - m2) // 2) * \
        f((J + M) // 2) * f((J - M) // 2)
    s = 0.0
    for k in range(0, a + 1):
        d = [k, a - k, (j1 - m1) // 2 - k, (j2 + m2) // 2 - k, (J - j2 + m1) // 2 + k, (J - j1 - m2) // 2 + k]
        if min(d) < 0:
            continue
        s += (-1) ** k / np.prod([float(f(x)) for x in d])
    return math.sqrt(pre) * s


class MBasis:
    """m-scheme single-particle states (orbit, 2m) of a J-scheme orbit table"""

    def __init__(self, jj, hole):
        self.jj = [int(x) for x in jj]
        self.states = [(o, m) for o in range(len(jj)) for m in range(-self.jj[o], self.jj[o] + 1, 2)]
        self.n = len(self.states)
        self.sl = {}
        k = 0
        for o in range(len(jj)):
            self.sl[o] = slice(k, k + self.jj[o] + 1)
            k += self.jj[o] + 1
        self.occ = np.array([float(hole[o]) for o, _ in self.states])
        self.ms = {o: list(range(-self.jj[o], self.jj[o] + 1, 2)) for o in range(len(jj))}

    # <ja ma jb mb | J M> as an array [ma, mb, M index (M = -J..J)]
    def _pair_cg(self, a, b, J):
        ma, mb = self.ms[a], self.ms[b]
        out = np.zeros((len(ma), len(mb), J + 1))
        for i, x in enumerate(ma):
            for k, y in enumerate(mb):
                M = x + y
                if abs(M) <= J:
                    out[i, k, (M + J) // 2] = cg(self.jj[a], x, self.jj[b], y, J, M)
        return out

    def _wigner(self, J1, J2, lam, mu):
        """<J2 M2 lam mu | J1 M1> / sqrt(J1 + 1) as [M2 index, M1 index]"""
        out = np.zeros((J2 + 1, J1 + 1))
        for i, M2 in enumerate(range(-J2, J2 + 1, 2)):
            M1 = M2 + mu
            if abs(M1) <= J1:
                out[i, (M1 + J1) // 2] = cg(J2, M2, lam, mu, J1, M1) / math.sqrt(J1 + 1)
        return out

    def one_body(self, elem, lam, mu):
        """t[(a ma), (b mb)] = <jb mb lam mu | ja ma> / sqrt(ja + 1) <a||t||b>"""
        t = np.zeros((self.n, self.n))
        for a in self.sl:
            for b in self.sl:
                r = elem(a, b)
                if r == 0.0:
                    continue
                for i, ma in enumerate(self.ms[a]):
                    for k, mb in enumerate(self.ms[b]):
                        c = cg(self.jj[b], mb, lam, mu, self.jj[a], ma)
                        if c:
                            t[self.sl[a].start + i, self.sl[b].start + k] = c / math.sqrt(self.jj[a] + 1) * r
        return t

    def two_body(self, elem, lam, mu):
        """T[(a)(b)(c)(d)] = sum_{J1 J2} CG(ab|J1) CG(cd|J2) <J2 lam|J1> / sqrt(J1+1)
        <ab J1||T||cd J2>, elem(a, b, c, d, J1, J2) the unnormalised reduced element"""
        T = np.zeros((self.n,) * 4)
        no = len(self.jj)
        for a in range(no):
            for b in range(no):
                for J1 in range(abs(self.jj[a] - self.jj[b]), self.jj[a] + self.jj[b] + 1, 2):
                    cab = self._pair_cg(a, b, J1)
                    for c in range(no):
                        for d in range(no):
                            for J2 in range(abs(self.jj[c] - self.jj[d]), self.jj[c] + self.jj[d] + 1, 2):
                                if J1 > J2 + lam or J2 > J1 + lam or J1 + J2 < lam:
                                    continue
                                r = elem(a, b, c, d, J1, J2)
                                if r == 0.0:
                                    continue
                                w = self._wigner(J1, J2, lam, mu)
                                ccd = self._pair_cg(c, d, J2)
                                T[self.sl[a], self.sl[b], self.sl[c], self.sl[d]] += r * np.einsum(
                                    "xyM,zwN,NM->xyzw", cab, ccd, w)
        return T

    def reduce_one(self, t, a, b, lam, mu):
        """<a||t||b> from the m-scheme matrix (projection on the Wigner-Eckart form)"""
        s = 0.0
        for i, ma in enumerate(self.ms[a]):
            for k, mb in enumerate(self.ms[b]):
                c = cg(self.jj[b], mb, lam, mu, self.jj[a], ma)
                if c:
                    s += c * t[self.sl[a].start + i, self.sl[b].start + k]
        return s * (lam + 1) / math.sqrt(self.jj[a] + 1)

    def reduce_two(self, T, a, b, c, d, J1, J2, lam, mu):
        """<ab J1||T||cd J2> (unnormalised) from the m-scheme array"""
        cab, ccd, w = self._pair_cg(a, b, J1), self._pair_cg(c, d, J2), self._wigner(J1, J2, lam, mu)
        blk = T[self.sl[a], self.sl[b], self.sl[c], self.sl[d]]
        s = np.einsum("xyzw,xyM,zwN,NM->", blk, cab, ccd, w)
        # sum over M1, M2 of w^2 (J1 + 1) = (J1 + 1) / (lam + 1)
        return s * (lam + 1) / (J1 + 1) * (J1 + 1)


def commutator(a1, a2, b1, b2, occ):
    """one- and two-body parts of [A, B] for normal-ordered A, B (IMSRG(2))"""
    n = occ
    nb = 1.0 - n
    D = n[:, None] - n[None, :]                     # n_a - n_b
    c1 = a1 @ b1 - b1 @ a1
    c1 += np.einsum("ab,biaj->ij", D * a1, b2) - np.einsum("ab,biaj->ij", D * b1, a2)
    W = n[:, None, None] * n[None, :, None] * nb[None, None, :] + nb[:, None, None] * nb[None, :, None] * n[None, None, :]
    c1 += 0.5 * (np.einsum("ciab,abc,abcj->ij", a2, W, b2, optimize=True) -
                 np.einsum("ciab,abc,abcj->ij", b2, W, a2, optimize=True))
    t1 = np.einsum("ia,ajkl->ijkl", a1, b2) - np.einsum("ia,ajkl->ijkl", b1, a2)
    t2 = np.einsum("ak,ijal->ijkl", a1, b2) - np.einsum("ak,ijal->ijkl", b1, a2)
    c2 = t1 - t1.transpose(1, 0, 2, 3) - (t2 - t2.transpose(0, 1, 3, 2))
    F = 1.0 - n[:, None] - n[None, :]
    c2 += 0.5 * (np.einsum("ijab,ab,abkl->ijkl", a2, F, b2, optimize=True) -
                 np.einsum("ijab,ab,abkl->ijkl", b2, F, a2, optimize=True))
    Y = np.einsum("ab,bjal,aibk->ijkl", D, a2, b2, optimize=True)
    c2 -= Y - Y.transpose(1, 0, 2, 3) - Y.transpose(0, 1, 3, 2) + Y.transpose(1, 0, 3, 2)
    return c1, c2
