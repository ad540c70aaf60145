"""A physical Hamiltonian: the harmonic-oscillator one-body part
(calculate_h0_harm_osc, basic_IMSRG.f90:2093-2191) and the me2j interaction
reader (read_me2j_interaction, interaction_IO.f90:282-557) with its
centre-of-mass two-body terms (r1_r2 / p1_p2, basic_IMSRG.f90:4445-4657).

The oracle's restatement is pinned by closed forms that share no formula
with it: HO radial matrix elements by quadrature of Laguerre functions, the
HO energies, the exact two-nucleon centre-of-mass spectrum (N_cm + 3/2) hw,
and the isospin Clebsch-Gordan decomposition of proton-neutron elements.
The GPU tests compare the library with the oracle on synthetic me2j files
(the reference's own file is a missing blob).
"""
import ctypes as C

import numpy as np
import pytest

from me2j_files import quadruples, synthetic_values, write_me2j

HW = 20.0


def _oracle(emax, holes=True):
    from eom_imsrg import hk_orbits, oxygen16_holes
    from oracle.oracle import OracleBasis
    o = hk_orbits(emax)
    h = oxygen16_holes(o) if holes else np.zeros(len(o["n"]), np.int32)
    return o, OracleBasis(o["n"], o["l"], o["jj"], o["tz"], h)


def _radial(n, l):
    """HO radial function (b = 1, positive at the origin) and its derivative"""
    from scipy.special import eval_genlaguerre, gammaln
    a = l + 0.5
    norm = np.sqrt(2.0 * np.exp(gammaln(n + 1) - gammaln(n + a + 1)))

    def R(r):
        return norm * r ** l * np.exp(-r * r / 2) * eval_genlaguerre(n, a, r * r)

    def dR(r):
        L = eval_genlaguerre(n, a, r * r)
        dL = -eval_genlaguerre(n - 1, a + 1, r * r) if n > 0 else 0.0
        lead = l * r ** (l - 1) if l > 0 else 0.0
        return norm * np.exp(-r * r / 2) * ((lead - r ** (l + 1)) * L + r ** l * 2 * r * dL)
    return R, dR


def _quad(f):
    from scipy.integrate import quad
    return quad(f, 0, 12, limit=400, epsabs=1e-13, epsrel=1e-13)[0]


def test_ho_one_body_vs_quadrature():
    """T (htype 3) and the r^2 part (htype 5, A = 1) against integrals of the
    radial functions: T = hw/2 <R_n'|-nabla^2|R_n>, U = hw/2 <R_n'|r^2|R_n>"""
    o, ob = _oracle(4)
    T = ob.h0_harm_osc(HW, 3, 16)
    U = ob.h0_harm_osc(HW, 5, 1)
    checked = 0
    for i in range(ob.N):
        for j in range(ob.N):
            if (o["l"][i], o["jj"][i], o["tz"][i]) != (o["l"][j], o["jj"][j], o["tz"][j]):
                continue
            l, ni, nj = int(o["l"][i]), int(o["n"][i]), int(o["n"][j])
            Ri, dRi = _radial(ni, l)
            Rj, dRj = _radial(nj, l)
            t = 0.5 * HW * _quad(lambda r: (dRi(r) * dRj(r) + l * (l + 1) * Ri(r) * Rj(r) / (r * r)) * r * r)
            u = 0.5 * HW * _quad(lambda r: Ri(r) * Rj(r) * r ** 4)
            # single-precision square roots in the off-diagonal elements
            tol = 1e-10 if ni == nj else 2e-7 * HW
            assert abs(ob.f_elem(T, 1, i, j) - t) < tol, (i, j, ob.f_elem(T, 1, i, j), t)
            assert abs(ob.f_elem(U, 1, i, j) - u) < tol, (i, j, ob.f_elem(U, 1, i, j), u)
            checked += 1
    assert checked > 50


def test_hamiltonian_types():
    """htype 2 = the HO energies hw (2n + l + 3/2) on the diagonal, zero
    elsewhere; htype 1 = T - T/A + beta (T + (com_hw/hw)^2 U)/A, i.e. T minus
    the one-body part of Tcm plus beta times that of Hcm"""
    o, ob = _oracle(4)
    h2 = ob.h0_harm_osc(HW, 2, 16)
    for i in range(ob.N):
        for j in range(ob.N):
            want = HW * (2 * o["n"][i] + o["l"][i] + 1.5) if i == j else 0.0
            assert abs(ob.f_elem(h2, 1, i, j) - want) < 1e-12
    A, beta, cm = 16, 3.0, 24.0
    h3, h4, h5 = ob.h0_harm_osc(HW, 3, A), ob.h0_harm_osc(HW, 4, A), ob.h0_harm_osc(HW, 5, A)
    h1 = ob.h0_harm_osc(HW, 1, A, beta, cm)
    np.testing.assert_allclose(h1, h3 - h4 + beta * h4 + beta * (cm / HW) ** 2 * h5, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(ob.h0_harm_osc(HW, 1, A), h3 - h4, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("emax", [2, 3])
def test_com_spectrum_pins_r1r2_p1p2(emax):
    """two distinguishable nucleons (proton a, neutron b): within the states of
    total quanta N = e_a + e_b at fixed J and parity,
        Hcm / hw = (N + 3)/2 + (r1.r2 + p1.p2)/2
    has the eigenvalues N_cm + 3/2, N_cm = 0..N (r1.r2 + p1.p2 = a1 a2^+ +
    a1^+ a2 conserves N).  A wrong sign, phase or 6j in either element
    leaves non-integer eigenvalues."""
    o, ob = _oracle(emax, holes=False)
    e = 2 * o["n"] + o["l"]
    prot = [i for i in range(ob.N) if o["tz"][i] == -1]
    neut = [i for i in range(ob.N) if o["tz"][i] == 1]
    nblocks = 0
    for N in range(emax + 1):
        for J in range(0, 2 * int(o["jj"].max()) + 1, 2):
            for P in (0, 1):
                st = [(a, b) for a in prot for b in neut if e[a] + e[b] == N and (o["l"][a] + o["l"][b]) % 2 == P
                      and abs(o["jj"][a] - o["jj"][b]) <= J <= o["jj"][a] + o["jj"][b]]
                if not st:
                    continue
                M = np.array([[0.5 * (ob.r1_r2(a, b, c, d, J) + ob.p1_p2(a, b, c, d, J)) for (c, d) in st]
                              for (a, b) in st])
                assert np.allclose(M, M.T, atol=1e-12)
                ev = np.linalg.eigvalsh(M + 0.5 * (N + 3) * np.eye(len(st))) - 1.5
                assert np.all(np.abs(ev - np.round(ev)) < 1e-10), (N, J, P, ev)
                assert ev.min() > -1e-10 and ev.max() < N + 1e-10
                nblocks += 1
                # the relative motion alone: r1.r2 - p1.p2 changes N by 2, so
                # it has no matrix elements inside one N
                R = np.array([[ob.r1_r2(a, b, c, d, J) - ob.p1_p2(a, b, c, d, J) for (c, d) in st]
                              for (a, b) in st])
                assert np.abs(R).max() < 1e-12
    assert nblocks > 5


def test_me2j_isospin_decomposition(tmp_path):
    """pp = T=1 (Mt +1 slot), nn = T=1 (Mt -1 slot), and for proton-neutron
    pairs the isospin Clebsch-Gordan forms <pn|V|pn> = (V1 + V0)/2,
    <pn|V|np> = (V1 - V0)/2 (orbits in the order written), read back through
    v_elem (phases and sqrt(2) of the stored pairs included)"""
    o, ob = _oracle(2)
    vals = synthetic_values(o, seed=3)
    # distinct T = 1 values per Mt, as a charge-dependent interaction has
    rng = np.random.default_rng(9)
    vals = vals.reshape(-1, 4)
    vals[:, 1] = np.where(vals[:, 2] != 0, vals[:, 2] + 0.1 * rng.standard_normal(len(vals)), 0.0)
    vals[:, 3] = np.where(vals[:, 2] != 0, vals[:, 2] - 0.1 * rng.standard_normal(len(vals)), 0.0)
    vals = np.round(vals.reshape(-1), 7)
    path = write_me2j(str(tmp_path / "v.me2j.gz"), vals)
    op = ob.read_me2j(path, o, HW, 3, 16)            # htype 3, beta 0: no centre-of-mass terms
    pidx = lambda k: 2 * k                            # proton entry of isospin orbit k
    nidx = lambda k: 2 * k + 1
    checked = 0
    for x, (k1, k2, kk1, kk2, J) in enumerate(quadruples(o)):
        v0, v_nn, v1, v_pp = vals[4 * x: 4 * x + 4]
        v = lambda a, b, c, d: ob.v_elem(op, 1, a, b, c, d, J)
        assert abs(v(pidx(k1), pidx(k2), pidx(kk1), pidx(kk2)) - v_pp) < 1e-12
        assert abs(v(nidx(k1), nidx(k2), nidx(kk1), nidx(kk2)) - v_nn) < 1e-12
        assert abs(v(pidx(k1), nidx(k2), pidx(kk1), nidx(kk2)) - 0.5 * (v1 + v0)) < 1e-12
        assert abs(v(pidx(k1), nidx(k2), nidx(kk1), pidx(kk2)) - 0.5 * (v1 - v0)) < 1e-12
        checked += 1
    assert checked == len(vals) // 4 and checked > 100


def test_me2j_parse_and_count(tmp_path):
    """the value count of the loop nest, and the fixed-width parse of both
    compressed and plain files"""
    from oracle.oracle import read_me2j_values
    o, ob = _oracle(2)
    vals = np.round(synthetic_values(o, seed=1), 7)
    assert len(vals) == 4 * sum(1 for _ in quadruples(o))
    for comp in (True, False):
        p = write_me2j(str(tmp_path / f"v{int(comp)}.me2j"), vals, compress=comp)
        np.testing.assert_array_equal(read_me2j_values(p, len(vals)), vals)


def test_read_me2j_errors_without_gpu(tmp_path):
    """file and orbit-table errors are reported before any device work"""
    from conftest import basis
    from eom_imsrg import _lib, hk_orbits
    lib = _lib.lib
    b = basis(2)
    o = hk_orbits(2)
    arr = [np.ascontiguousarray(o[k], np.int32) for k in ("n", "l", "jj", "tz")]
    ptr = [a.ctypes.data_as(C.POINTER(C.c_int)) for a in arr]
    x = C.c_void_p(16)   # never dereferenced: the checks fail first
    rc = lib.imsrg_read_me2j(b.handle, str(tmp_path / "missing.me2j.gz").encode(), len(arr[0]), *ptr, HW, 1, 0.0,
                             HW, 16, x, None)
    assert rc == _lib.IMSRG_EINVAL and b"cannot open" in lib.imsrg_last_error()
    short = write_me2j(str(tmp_path / "short.me2j.gz"), np.zeros(25))
    rc = lib.imsrg_read_me2j(b.handle, short.encode(), len(arr[0]), *ptr, HW, 1, 0.0, HW, 16, x, None)
    assert rc == _lib.IMSRG_EINVAL and b"ends after" in lib.imsrg_last_error()
    rc = lib.imsrg_read_me2j(b.handle, short.encode(), len(arr[0]) - 1, *ptr, HW, 1, 0.0, HW, 16, x, None)
    assert rc == _lib.IMSRG_EINVAL and b"even" in lib.imsrg_last_error()
    tz = arr[3].copy()
    tz[0], tz[1] = 1, -1
    ptr2 = ptr[:3] + [tz.ctypes.data_as(C.POINTER(C.c_int))]
    rc = lib.imsrg_read_me2j(b.handle, short.encode(), len(arr[0]), *ptr2, HW, 1, 0.0, HW, 16, x, None)
    assert rc == _lib.IMSRG_EINVAL and b"proton, neutron" in lib.imsrg_last_error()
    assert lib.imsrg_read_me2j(b.handle, short.encode(), len(arr[0]), *ptr, HW, 9, 0.0, HW, 16, x,
                               None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_h0_harm_osc(b.handle, HW, 7, 0.0, HW, 16, x, None) == _lib.IMSRG_EINVAL
    assert lib.imsrg_h0_harm_osc(None, HW, 1, 0.0, HW, 16, x, None) == _lib.IMSRG_EINVAL


# ---------------------------------------------------------------- GPU parity
def _gpu_basis(emax):
    from eom_imsrg import Basis, hk_orbits, oxygen16_holes
    from oracle.oracle import OracleBasis
    o = hk_orbits(emax)
    h = oxygen16_holes(o)
    return Basis(o["n"], o["l"], o["jj"], o["tz"], h, o["e"]), OracleBasis(o["n"], o["l"], o["jj"], o["tz"], h)


@pytest.mark.gpu
def test_h0_matches_oracle(gpu):
    """every Hamiltonian type, with and without the Lawson term: bitwise equal
    one-body parts; E0 and the two-body part untouched"""
    from eom_imsrg import SqOp, interaction
    b, ob = _gpu_basis(4)
    print("[h0] emax=4", flush=True)
    rng = np.random.default_rng(0)
    base = rng.standard_normal(b.oplen)
    for htype, beta, cm in ((1, 0.0, HW), (1, 3.0, 24.0), (2, 0.0, HW), (3, 0.0, HW), (4, 0.0, HW), (5, 0.0, HW)):
        H = SqOp.from_numpy(b, base.copy(), 1)
        interaction.h0_harm_osc(H, HW, htype, 16, beta, cm)
        ref = ob.h0_harm_osc(HW, htype, 16, beta, cm, op=base.copy())
        assert np.array_equal(H.numpy(), ref), (htype, beta)


@pytest.mark.gpu
@pytest.mark.parametrize("emax,file_emax", [(4, 4), (2, 4)])
@pytest.mark.parametrize("htype,beta", [(1, 0.0), (1, 3.0), (3, 0.0)])
def test_read_me2j_matches_oracle(gpu, tmp_path, emax, file_emax, htype, beta):
    """library reader vs oracle on the same synthetic file (element by
    element, 1e-11: the two 6j evaluations differ in the last bits); file
    orbits beyond the basis are skipped; elements the file does not hold
    keep their values"""
    from eom_imsrg import SqOp, hk_orbits, interaction
    b, ob = _gpu_basis(emax)
    fo = hk_orbits(file_emax)
    print(f"[me2j] emax={emax} file_emax={file_emax} htype={htype} beta={beta}", flush=True)
    path = write_me2j(str(tmp_path / "v.me2j.gz"), synthetic_values(fo, seed=emax + 10 * file_emax))
    base = np.full(b.oplen, 0.25)
    H = SqOp.from_numpy(b, base.copy(), 1)
    interaction.read_me2j(H, path, file_emax, HW, htype, 16, beta, 24.0)
    ref = ob.read_me2j(path, fo, HW, htype, 16, beta, 24.0, op=base.copy())
    got = H.numpy()
    assert np.array_equal(got == 0.25, ref == 0.25)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12)


@pytest.mark.gpu
def test_flow_from_me2j(gpu, tmp_path):
    """the reference's set-up order (read_me2j, calculate_h0_harm_osc,
    normal_order) and magnus_decouple on the GPU vs the oracle: equal steps,
    E0 per step within 1e-8 MeV"""
    from eom_imsrg import Flow, SqOp, commutators as cm, hk_orbits, interaction
    REF = float(np.float32(1e-6))
    b, ob = _gpu_basis(4)
    fo = hk_orbits(4)
    print("[me2j flow] emax=4", flush=True)
    path = write_me2j(str(tmp_path / "v.me2j.gz"), synthetic_values(fo, seed=5))
    H = SqOp.zeros(b, 1)
    interaction.read_me2j(H, path, 4, HW, 1, 16)
    interaction.h0_harm_osc(H, HW, 1, 16)
    cm.normal_order(H)
    Hr = ob.read_me2j(path, fo, HW, 1, 16)
    Hr = ob.h0_harm_osc(HW, 1, 16, op=Hr)
    Hr = ob.normal_order(Hr)
    np.testing.assert_allclose(H.numpy(), Hr, rtol=1e-11, atol=1e-11)
    HSr, _, logr, str_ = ob.magnus_decouple(Hr, 0.5, REF, REF, 0, REF)
    log = Flow(b, graphs=True).magnus_decouple(H, SqOp.zeros(b, -1))
    assert len(log) == str_
    assert np.array_equal(log[:, 3], logr[:, 3])
    assert np.max(np.abs(log[:, 1] - logr[:, 1])) < 1e-8
