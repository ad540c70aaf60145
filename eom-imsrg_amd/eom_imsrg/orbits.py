"""Harmonic-oscillator orbit tables in the reference's ``hk*.sps`` order.

The reference reads ``sp_inputs/hk<N>.sps`` (label n l 2j 2tz e) and keeps
the orbits with 2n+l <= eMax and l <= lMax (basic_IMSRG.f90:192-332).  Those
files list, shell by shell (e = 2n+l ascending), l ascending, then 2j
ascending, then 2tz = -1 (proton) before +1 (neutron).  ``hk_orbits`` builds
the same table from that rule; the test suite checks it against the
reference files' contents (tests/golden/hk_orbits.json).

Holes: ``sp_inputs/hole_scheme_hk`` marks occupied orbits per (Z, N); the
configurations here use 16O (Z = N = 8): the 0s1/2, 0p1/2 and 0p3/2 shells of
both isospins, i.e. the first six orbits.
"""
from __future__ import annotations

import numpy as np


def hk_orbits(emax: int, lmax: int | None = None) -> dict:
    """Orbit table (n, l, 2j, 2tz, e) for 2n+l <= emax, l <= lmax."""
    if lmax is None:
        lmax = emax
    n, l, jj, tz, e = [], [], [], [], []
    for shell in range(emax + 1):
        for ll in range(shell % 2, shell + 1, 2):
            if ll > lmax:
                continue
            nn = (shell - ll) // 2
            for j2 in (2 * ll - 1, 2 * ll + 1):
                if j2 < 1:
                    continue
                for t in (-1, 1):
                    n.append(nn); l.append(ll); jj.append(j2); tz.append(t); e.append(shell)
    return dict(n=np.array(n, np.int32), l=np.array(l, np.int32), jj=np.array(jj, np.int32),
                tz=np.array(tz, np.int32), e=np.array(e, np.int32))


def oxygen16_holes(orb: dict) -> np.ndarray:
    """Hole flags of the Z = N = 8 reference (row "8 8" of hole_scheme_hk)."""
    hole = np.zeros(len(orb["n"]), np.int32)
    for i in range(len(hole)):
        if orb["n"][i] == 0 and orb["l"][i] <= 1:
            hole[i] = 1
    return hole
