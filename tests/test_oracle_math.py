"""Closed-form and invariant checks of the C restatement: each pins a part
of the algorithm independently of the reference run (a dropped term, wrong
sign or wrong index fails one of them)."""
import ctypes as ct

import numpy as np
import pytest

from common import load_golden, oracle_available

pytestmark = pytest.mark.skipif(not oracle_available(), reason="oracle not built (oracle/Makefile)")


def lib(prec="dp"):
    import oracle
    L = oracle.load_lib(prec)
    L.fwo_hll.restype = ct.c_double
    L.fwo_hll.argtypes = [ct.c_double] * 6
    L.fwo_thomas_line.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    L.fwo_muscl_row.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    return L


def test_hll_branches_and_closed_form():
    L = lib()
    rng = np.random.default_rng(0)
    for _ in range(2000):
        sl, sr = np.sort(rng.normal(size=2) * 3)
        fl, fr, ul, ur = rng.normal(size=4)
        got = L.fwo_hll(sl, sr, fl, fr, ul, ur)
        if sl >= 0:
            exp = fl
        elif sr <= 0:
            exp = fr
        else:
            den = sr - sl if abs(sr - sl) >= 1e-6 else 1e-6
            exp = (sr * fl - sl * fr + sl * sr * (ur - ul)) / den
        assert got == pytest.approx(exp, rel=1e-14, abs=1e-14)
    # consistency: equal states give the physical flux
    assert L.fwo_hll(-1.0, 2.0, 0.7, 0.7, 0.3, 0.3) == pytest.approx(0.7, rel=1e-15)


def dense_solve(a, c, d):
    n = len(d)
    A = np.eye(n)
    for k in range(n):
        if k > 0:
            A[k, k - 1] = a[k]
        if k < n - 1:
            A[k, k + 1] = c[k]
    return np.linalg.solve(A, d)


@pytest.mark.parametrize("n", [2, 5, 64, 300])
def test_thomas_matches_dense_solve(n):
    L = lib()
    rng = np.random.default_rng(n)
    a = rng.uniform(-0.3, 0.3, n)
    c = rng.uniform(-0.3, 0.3, n)
    d = rng.normal(size=n)
    a[rng.uniform(size=n) < 0.2] = 0.0  # rows skipped by the sweep (a == 0)
    a[0] = 0.0
    c[-1] = 0.0
    exp = dense_solve(a, c, d)
    z = np.zeros(n)
    aa, cc, dd = a.copy(), c.copy(), d.copy()
    L.fwo_thomas_line(n, aa.ctypes.data, cc.ctypes.data, dd.ctypes.data, z.ctypes.data)
    np.testing.assert_allclose(z, exp, rtol=1e-12, atol=1e-12)


def test_muscl_reproduces_linear_data_and_is_tvd():
    L = lib()
    n = 40
    x = np.arange(n, dtype=np.float64)
    v = 0.25 + 0.5 * x
    mk = np.ones(n, np.int32)
    ol = np.zeros(n + 1)
    orr = np.zeros(n + 1)
    L.fwo_muscl_row(n, v.ctypes.data, mk.ctypes.data, ol.ctypes.data, orr.ctypes.data)
    faces = np.arange(4, n - 2)  # 1-based faces Ibeg..Iend+1
    # face i (1-based) sits between cells i-1 and i -> x = i - 1.5 (0-based cells)
    exact = 0.25 + 0.5 * (faces - 1.5)
    np.testing.assert_allclose(ol[faces - 1], exact, rtol=0, atol=1e-12)
    np.testing.assert_allclose(orr[faces - 1], exact, rtol=0, atol=1e-12)
    # monotone random data: reconstructed states stay within neighbour bounds
    rng = np.random.default_rng(3)
    v = np.cumsum(rng.uniform(0, 1, n))
    L.fwo_muscl_row(n, v.ctypes.data, mk.ctypes.data, ol.ctypes.data, orr.ctypes.data)
    for f in faces:
        lo, hi = v[f - 2], v[f - 1]  # cells left/right of face f (0-based f-2, f-1)
        assert lo - 1e-12 <= ol[f - 1] <= hi + 1e-12
        assert lo - 1e-12 <= orr[f - 1] <= hi + 1e-12


def _flat_state(prec, eta_fn, m=48, n=40, depth=5.0, dispersion=True):
    """A closed flat-bottom basin built on top of a fixture's scalars."""
    init, _ = load_golden("island", prec)
    rdt = np.float64 if prec == "dp" else np.float32
    st = {k: v for k, v in init.items() if not isinstance(v, np.ndarray)}
    M, N = m + 6, n + 6
    st.update(Mglob=m, Nglob=n, Mloc=M, Nloc=N, Ibeg=4, Iend=M - 3, Jbeg=4, Jend=N - 3,
              WaveMaker="INI_SOL", DIRECT_SPONGE=0, PERIODIC=0, TIME=0.0, DT=0.0,
              DISPERSION=int(dispersion), Gamma1=1.0 if dispersion else 0.0,
              Gamma2=1.0 if dispersion else 0.0, EtaBlowVal=1e9)
    yy, xx = np.meshgrid(np.arange(N) - 2.5 - n / 2, np.arange(M) - 2.5 - m / 2, indexing="ij")
    st["Depth"] = np.full((N, M), depth, rdt)
    st["DepthX"] = np.full((N, M + 1), depth, rdt)
    st["DepthY"] = np.full((N + 1, M), depth, rdt)
    eta = eta_fn(xx, yy).astype(rdt)
    st["Eta"] = eta
    st["H"] = (eta + depth).astype(rdt)
    st["MASK"] = np.ones((N, M), np.int32)
    st["MASK9"] = np.ones((N, M), np.int32)
    st["MASK_STRUC"] = np.ones((N, M), np.int32)
    st["SPONGE"] = np.ones((N, M), rdt)
    st["SP_BYTES"] = 8 if prec == "dp" else 4
    return st


def test_lake_at_rest_stays_at_rest_to_small():
    """Well-balanced pressure split (fluxes.F:1365-1373): still water over
    arbitrary bathymetry stays still.  Not exactly: the van Leer clamp turns a
    zero slope into +-SMALL (fluxes.F:1609-1610), so a flat field is
    reconstructed as +-SMALL/2 = 5e-7 at faces; the drift stays at that level
    and must not grow."""
    import oracle
    st = _flat_state("dp", lambda x, y: 0.0 * x)
    rng = np.random.default_rng(7)
    N, M = st["Depth"].shape
    bump = rng.uniform(2.0, 6.0, (N, M))
    st["Depth"] = bump
    st["H"] = bump.copy()
    st["DepthX"] = np.zeros((N, M + 1))
    st["DepthX"][:, 1:M] = 0.5 * (bump[:, :-1] + bump[:, 1:])
    st["DepthX"][:, 0] = bump[:, 0]
    st["DepthX"][:, M] = bump[:, -1]
    st["DepthY"] = np.zeros((N + 1, M))
    st["DepthY"][1:N, :] = 0.5 * (bump[:-1, :] + bump[1:, :])
    st["DepthY"][0, :] = bump[0, :]
    st["DepthY"][N, :] = bump[-1, :]
    o = oracle.OracleSolver(st)
    assert o.run(50) == 0
    e50 = np.abs(o["Eta"][3:-3, 3:-3]).max()
    assert o.run(150) == 0
    e200 = np.abs(o["Eta"][3:-3, 3:-3]).max()
    assert e50 < 2e-5 and e200 < 2e-5, (e50, e200)
    for k in ("U", "V"):
        assert np.abs(o[k][3:-3, 3:-3]).max() < 1e-4, k


def test_mass_conserved_in_closed_basin():
    import oracle
    st = _flat_state("dp", lambda x, y: 0.1 * np.exp(-(x ** 2 + y ** 2) / 30.0))
    o = oracle.OracleSolver(st)
    m0 = o["Eta"][3:-3, 3:-3].sum()
    assert o.run(150) == 0
    m1 = o["Eta"][3:-3, 3:-3].sum()
    assert abs(m1 - m0) <= 1e-11 * max(1.0, abs(m0))
    assert np.abs(o["U"]).max() > 1e-4  # it did move


def test_mirror_symmetry_of_symmetric_hump():
    """x -> -x, y -> -y and x <-> y symmetry of a symmetric hump on a flat
    bottom.  The SMALL clamp biases left states up and right states down
    (fluxes.F:1609-1635), which is not mirror symmetric, so symmetry holds to
    a few SMALL (1e-6), far below the hump amplitude (5e-2); a sign or index
    error breaks it at the amplitude level."""
    import oracle
    st = _flat_state("dp", lambda x, y: 0.05 * np.exp(-(x ** 2 + y ** 2) / 20.0), m=40, n=40)
    o = oracle.OracleSolver(st)
    assert o.run(80) == 0
    e = o["Eta"][3:-3, 3:-3]
    u = o["U"][3:-3, 3:-3]
    v = o["V"][3:-3, 3:-3]
    tol = 1e-5
    assert np.abs(e).max() > 1e-3
    np.testing.assert_allclose(e, e[:, ::-1], atol=tol)      # x -> -x
    np.testing.assert_allclose(e, e[::-1, :], atol=tol)      # y -> -y
    np.testing.assert_allclose(u, -u[:, ::-1], atol=tol)     # u odd in x
    np.testing.assert_allclose(v, -v[::-1, :], atol=tol)     # v odd in y
    np.testing.assert_allclose(e, e.T, atol=tol)             # x <-> y


@pytest.mark.parametrize("dispersion", [True, False])
def test_standing_wave_period_follows_dispersion_relation(dispersion):
    """Linear standing wave in a closed flat basin: the period must follow
    the model's linear dispersion relation, omega^2 = g h k^2 (1 - (a + 1/3)
    (kh)^2) / (1 - a (kh)^2) with a = beta^2/2 + beta, beta = -0.531 (Nwogu /
    Kennedy form), or the shallow-water one without dispersion."""
    import oracle
    h, m, amp = 4.0, 128, 1e-3
    L = float(m)                       # basin length (dx = 1 m in the fixture)
    init, _ = load_golden("island", "dp")
    dx = init["DX"]
    L = m * dx
    k = 2 * np.pi / L * 2              # mode with two wavelengths in the basin
    st = _flat_state("dp", lambda x, y: amp * np.cos(k * (x + m / 2 + 0.5) * dx) + 0 * y,
                     m=m, n=4, depth=h, dispersion=dispersion)
    beta = -0.531
    a = beta ** 2 / 2 + beta
    kh = k * h
    if dispersion:
        w2 = 9.81 * h * k * k * (1 - (a + 1 / 3) * kh ** 2) / (1 - a * kh ** 2)
    else:
        w2 = 9.81 * h * k * k
    period = 2 * np.pi / np.sqrt(w2)
    o = oracle.OracleSolver(st)
    ts, es = [], []
    while o.time.value < 1.6 * period:
        o.run(1)
        ts.append(o.time.value)
        es.append(o["Eta"][5, 3])      # antinode at the west wall
    ts, es = np.array(ts), np.array(es)
    # time of the first minimum -> half period
    i = np.argmin(es[: len(es) * 3 // 4])
    half = ts[i]
    assert half == pytest.approx(period / 2, rel=0.02)
