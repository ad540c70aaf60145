"""Synthetic me2j interaction files for the tests (the reference ships none:
its chi2b_srg0625_eMax08_hwHO020.me2j.gz is listed as a missing blob).

Layout of the file (interaction_IO.f90:318-392): one header line, then the
values in the loop order of the isospin-coupled orbit quadruples (k1 >= k2,
kk1 <= k1, kk2 <= kk1 or <= k2 when kk1 == k1, equal parity, every common
2J), four per (quadruple, 2J) in the order (T, Mt) = (0, 0), (1, -1),
(1, 0), (1, 1), written ten a line as f12.7 then f13.7.
"""
from __future__ import annotations

import gzip

import numpy as np


def quadruples(orbits):
    """(k1, k2, kk1, kk2, 2J) in file order; orbits = the file's pn table"""
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


def synthetic_values(orbits, seed=0, scale=2.0):
    """N(0,1) * scale * exp(-(e1+e2+e3+e4)/8) per isospin component, with the
    components a pair of identical isospin orbits cannot carry set to zero
    (two nucleons in one j-shell: 2J/2 even needs T = 1, odd needs T = 0),
    and charge-independent T = 1 values (same for every Mt)"""
    rng = np.random.default_rng(seed)
    e = 2 * np.asarray(orbits["n"])[::2] + np.asarray(orbits["l"])[::2]
    out = []
    for k1, k2, kk1, kk2, J in quadruples(orbits):
        w = scale * np.exp(-(e[k1] + e[k2] + e[kk1] + e[kk2]) / 8.0)
        v0, v1 = rng.standard_normal(2) * w
        ident = k1 == k2 or kk1 == kk2
        if ident and (J // 2) % 2 == 0:
            v0 = 0.0
        if ident and (J // 2) % 2 == 1:
            v1 = 0.0
        out.extend((v0, v1, v1, v1))
    return np.array(out)


def write_me2j(path, values, compress=True):
    """the values as an me2j file (text, gzip-compressed unless told not to)"""
    v = np.round(np.asarray(values, np.float64), 7)
    lines = ["synthetic me2j interaction file\n"]
    for i in range(0, len(v), 10):
        chunk = v[i: i + 10]
        lines.append(f"{chunk[0]:12.7f}" + "".join(f"{x:13.7f}" for x in chunk[1:]) + "\n")
    data = "".join(lines).encode()
    if compress:
        with gzip.open(path, "wb") as f:
            f.write(data)
    else:
        with open(path, "wb") as f:
            f.write(data)
    return path
