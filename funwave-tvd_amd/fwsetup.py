"""Host-side initial state for the per-step solver, from a workload's input
parameters (fwcases.Case), without running the reference program.

Restates the parts of the reference initialisation that produce the arrays
and scalars the per-step path reads (src/init.F INITIALIZATION, src/io.F
READ_INPUT defaults, src/sponge.F CALCULATE_SPONGE, src/wavemaker.F
WK_REG / WK_DATA2D set-up).  It is evaluated in float64 with numpy and
rounded once to the working precision, so it agrees with the reference's
own initialisation to rounding (checked in tests/test_setup.py), not bit for
bit; the per-step parity tests always start both sides from one state.
"""
from __future__ import annotations

import numpy as np

GRAV = 9.81
PI_SP = float(np.float32(3.141592653))  # default-REAL pi, src/mod_param.F:71
SMALL = 1e-6
LARGE = 999999.0
DEG2RAD = 0.0175                        # src/mod_param.F:79


def _phi_coll(a, vtype, periodic, ib, ie, jb, je):
    """Ghost fill of one cell array (bc.F:631-1272), [j, i] indexing."""
    sx = -1.0 if vtype == 2 else 1.0
    sy = -1.0 if vtype == 3 else 1.0
    M = a.shape[1]
    if periodic:
        for k in range(1, 4):
            a[k - 1, :] = a[je - 3 + k - 1, :]
            a[je + k - 1, :] = a[3 + k - 1, :]
    for k in range(1, 4):
        a[jb - 1:je, k - 1] = sx * a[jb - 1:je, ib + 3 - k - 1]
        a[jb - 1:je, ie + k - 1] = sx * a[jb - 1:je, ie - k + 1 - 1]
    if not periodic:
        for k in range(1, 4):
            a[k - 1, :M] = sy * a[jb + 3 - k - 1, :M]
            a[je + k - 1, :M] = sy * a[je - k + 1 - 1, :M]


def _wkn(omega, h, alpha=-0.39):
    """Wave number of the extended Boussinesq dispersion relation used by
    the Wei-Kirby source (wavemaker.F:2380-2393)."""
    alpha1 = alpha + 1.0 / 3.0
    tb = omega * omega * h / GRAV
    tc = 1.0 + tb * alpha
    return np.sqrt((tc - np.sqrt(tc * tc - 4.0 * alpha1 * tb)) / (2.0 * alpha1)) / h


def _wk_source_coeffs(freq, amp, theta, h, delta, alpha=-0.39):
    """D_gen, beta_gen, rlamda of one component (wavemaker.F:2369-2405,
    :779-800)."""
    alpha1 = alpha + 1.0 / 3.0
    omgn = 2.0 * PI_SP * freq
    wkn = _wkn(omgn, h, alpha)
    c_phase = 1.0 / wkn * freq * 2.0 * PI_SP
    wave_length = c_phase / freq
    rlamda = wkn * np.sin(theta)
    beta_gen = 80.0 / delta ** 2 / wave_length ** 2
    rl_gen = wkn * np.cos(theta)
    ri = np.sqrt(3.14159 / beta_gen) * np.exp(-rl_gen ** 2 / 4.0 / beta_gen)
    d_gen = 2.0 * amp * np.cos(theta) * (omgn ** 2 - alpha1 * GRAV * wkn ** 4 * h ** 3) / (
        omgn * wkn * ri * (1.0 - alpha * (wkn * h) ** 2))
    return d_gen, beta_gen, rlamda, wave_length


def _sponge(M, N, dx, dy, widths, r_sp, a_sp):
    """CALCULATE_SPONGE (sponge.F:591-767), serial branch."""
    sp = np.ones((N, M))
    I = np.arange(1, M + 1)[None, :]
    J = np.arange(1, N + 1)[:, None]
    w, e, s, n = widths
    t1 = np.zeros((N, M))
    t2 = np.zeros((N, M))
    floor = a_sp ** (r_sp ** 50)

    def lim_of(a):
        return np.where(a > 1.0, a, 1.0)

    if w > 0:
        iw = int(w / dx) + 3
        ri = r_sp ** ((50 * (I - 1)) // (iw - 1))
        t1 = np.maximum(a_sp ** ri, lim_of(sp)) * np.ones((N, 1))
    if e > 0:
        iw = int(e / dx) + 3
        ri = r_sp ** ((50 * (M - I)) // (iw - 1))
        t2 = np.maximum(a_sp ** ri, lim_of(sp)) * np.ones((N, 1))
    sp = np.where(t1 > t2, t1, t2)
    sp = np.where(sp < floor, 1.0, sp)
    # tmp_2d_1 / tmp_2d_2 keep the x-pass values when a y side has no sponge
    if s > 0:
        iw = int(s / dy) + 3
        ri = r_sp ** ((50 * (J - 1)) // (iw - 1))
        t1 = np.maximum(a_sp ** ri, lim_of(sp))
    if n > 0:
        iw = int(n / dy) + 3
        ri = r_sp ** ((50 * (N - J)) // (iw - 1))
        t2 = np.maximum(a_sp ** ri, lim_of(sp))
    sp = np.where(t1 > t2, t1, t2)
    sp = np.where(sp < floor, 1.0, sp)
    return sp


def _solitary_coeffs(a0, h1, rdt):
    """SUB_SLTRY (samples.F:582-634) in the working precision: phase speed and
    the coefficients of u = au sech^2, eta = ae1 sech^2 + ae2 sech^4.  alpha
    is the default-REAL literal -0.39 of INITIAL_SOLITARY_WAVE (:531)."""
    one, two, three = rdt(1), rdt(2), rdt(3)
    a0, h1 = rdt(a0), rdt(h1)
    alp = rdt(np.float32(-0.39))
    alp2 = alp + one / three
    eps = a0 / h1
    p = -(alp2 + two * alp * (one + eps)) / (two * alp)
    q = eps * alp2 / alp
    r = alp2 / (two * alp)
    x = rdt(1.2)
    for _ in range(11):  # Newton for the root x > 1 of x^3 + p x^2 + q x + r
        fx = r + x * (q + x * (p + x))
        fpx = q + x * (two * p + three * x)
        x = x - fx / fpx
        if not abs(fx) >= rdt(np.float32(1e-5)):
            break
    else:
        raise ValueError("no solitary wave solution for AMP/DEP = %g" % (a0 / h1))
    rx = np.sqrt(x)
    cph = np.sqrt(rdt(GRAV) * h1)
    au = (x - one) / (eps * rx) * cph * eps
    bue = np.sqrt((x - one) / (rdt(4) * (alp2 - alp * x))) / h1
    ae1 = (x - one) / (eps * three * (alp2 - alp * x)) * a0
    ae2 = -((x - one) / (two * eps) * (x - one) * (two * alp * x + alp2) / (x * (alp2 - alp * x)) * a0)
    return rx * cph, bue, ae1, ae2, au


def _tanh_sponge(M, N, dx, dy, widths, coef):
    """CALCULATE_FRICTION_SPONGE / CALCULATE_DIFFUSION_SPONGE
    (sponge.F:223-395, :405-580) from a zero field: coef * tanh(ri / 10) with
    ri the number of cells inside the layer (ghosts included), west/east
    merged by max, then south/north limited below by the merged field.  As
    in the reference the two temporaries are not cleared between the passes."""
    west, east, south, north = widths
    sp = np.zeros((N, M))
    t1 = np.zeros((N, M))
    t2 = np.zeros((N, M))
    I = np.arange(1, M + 1)[None, :]
    J = np.arange(1, N + 1)[:, None]
    if west > 0:
        iw = int(west / dx) + 3
        t1 = np.broadcast_to(np.maximum(coef * np.tanh(np.maximum(0.0, iw - I) / 10.0), 0.0), (N, M)).copy()
    if east > 0:
        iw = int(east / dx) + 3
        t2 = np.broadcast_to(np.maximum(coef * np.tanh(np.maximum(0.0, iw - M + I) / 10.0), 0.0), (N, M)).copy()
    sp = np.where(t1 > t2, t1, t2)
    lim = np.where(sp > 0.0, sp, 0.0)
    if south > 0:
        iw = int(south / dy) + 3
        t1 = np.maximum(coef * np.tanh(np.maximum(0.0, iw - J) / 10.0), lim)
    if north > 0:
        iw = int(north / dy) + 3
        t2 = np.maximum(coef * np.tanh(np.maximum(0.0, iw - N + J) / 10.0), lim)
    return np.where(t1 > t2, t1, t2)


def build_state(case, precision: str = "sp") -> dict:
    """Initial state (reference names, arrays indexed [j, i]) of a case."""
    rdt = np.float64 if precision == "dp" else np.float32
    p = case.params
    mg, ng = int(p["Mglob"]), int(p["Nglob"])
    M, N = mg + 6, ng + 6
    ib, ie, jb, je = 4, M - 3, 4, N - 3
    dx, dy = float(p["DX"]), float(p["DY"])
    periodic = bool(p.get("PERIODIC", False))
    disp = bool(p.get("DISPERSION", False))
    mindep = min(float(p.get("MinDepth", 0.1)), float(p.get("MinDepthFrc", 0.1)))  # io.F:2930-2931
    st = {"SP_BYTES": 8 if precision == "dp" else 4,
          "Mglob": mg, "Nglob": ng, "Mloc": M, "Nloc": N, "Ibeg": ib, "Iend": ie,
          "Jbeg": jb, "Jend": je, "Nghost": 3, "DX": dx, "DY": dy, "DXg": dx, "DYg": dy,
          "DT": 0.0, "TIME": 0.0, "CFL": float(p.get("CFL", 0.5)),
          "MinDepth": mindep, "MinDepthFrc": mindep,
          "DISPERSION": int(disp), "PERIODIC": int(periodic)}
    beta_ref = float(p.get("Beta_ref", -0.531))
    g1 = float(p.get("Gamma1", 1.0)) if disp else 0.0
    g2 = float(p.get("Gamma2", 1.0)) if disp else 0.0
    st.update(Gamma1=g1, Gamma2=g2, Gamma3=float(p.get("Gamma3", 1.0)),
              Beta_1=beta_ref + 1.0, Beta_2=(1.0 / 5.0) ** 2, b1=beta_ref * beta_ref, b2=beta_ref,
              SWE_ETA_DEP=float(p.get("SWE_ETA_DEP", 0.8)), FroudeCap=float(p.get("FroudeCap", 3.0)),
              Cbrk1=float(p.get("Cbrk1", 0.65)), Cbrk2=float(p.get("Cbrk2", 0.35)),
              WAVEMAKER_Cbrk=float(p.get("WAVEMAKER_Cbrk", 1.0)), nu_bkg=float(p.get("nu_bkg", 0.0)),
              VISCOSITY_BREAKING=int(bool(p.get("VISCOSITY_BREAKING", False))),
              SHOW_BREAKING=int(bool(p.get("SHOW_BREAKING", True))),
              DIFFUSION_SPONGE=int(bool(p.get("DIFFUSION_SPONGE", False))),
              FRICTION_SPONGE=int(bool(p.get("FRICTION_SPONGE", False))),
              DIRECT_SPONGE=int(bool(p.get("DIRECT_SPONGE", False))),
              # a DT_fixed entry switches FIXED_DT on (io.F:2882-2896)
              ETA_LIMITER=0, FIXED_DT=int("DT_fixed" in p), DT_fixed=float(p.get("DT_fixed", 0.0)),
              CrestLimit=0.0, TroughLimit=0.0,
              T_brk=20.0, WaveMaker=str(p.get("WAVEMAKER", "NONE")))

    # bathymetry, src/init.F:684-777, then mirror ghosts and face depths :789-829
    depth = np.full((N, M), 10.0)  # default before the bathymetry is read (init.F:548)
    dtype = str(p.get("DEPTH_TYPE", "FLAT")).upper()
    if dtype.startswith("DAT"):
        depth[jb - 1:je, ib - 1:ie] = np.asarray(case.depth, np.float64)
    elif dtype.startswith("FLA"):
        depth[:, :] = float(p["DEPTH_FLAT"])
    elif dtype.startswith("SLO"):
        df, slp, xs = float(p["DEPTH_FLAT"]), float(p["SLP"]), float(p["Xslp"])
        depth[jb - 1:je, ib - 1:ie] = df
        i0 = int(xs / dx) + 1 + ib
        for i in range(i0, ie + 1):
            depth[jb - 1:je, i - 1] = df - slp * (i - i0) * dx
    else:
        raise ValueError(dtype)
    _phi_coll(depth, 1, periodic, ib, ie, jb, je)
    depx = np.zeros((N, M + 1))
    depx[:, 1:M] = 0.5 * (depth[:, :-1] + depth[:, 1:])
    depx[:, 0] = 0.5 * (3.0 * depth[:, 0] - depth[:, 1])
    depx[:, M] = 0.5 * (3.0 * depth[:, M - 1] - depth[:, M - 2])
    depy = np.zeros((N + 1, M))
    depy[1:N, :] = 0.5 * (depth[:-1, :] + depth[1:, :])
    depy[0, :] = 0.5 * (3.0 * depth[0, :] - depth[1, :])
    depy[N, :] = 0.5 * (3.0 * depth[N - 1, :] - depth[N - 2, :])
    st["EtaBlowVal"] = 100.0 * np.abs(depth[jb - 1:je, ib - 1:ie]).max()  # init.F:847
    st["Cd"] = np.full((N, M), float(p.get("Cd", 0.0)))                     # init.F:883-889

    # wavemaker, src/wavemaker.F:55-613 and the band of init.F:1236-1268
    wk = st["WaveMaker"].strip()
    delta = float(p.get("Delta_WK", 0.5))
    st.update(Xc_WK=float(p.get("Xc_WK", 0.0)), Yc_WK=float(p.get("Yc_WK", 0.0)),
              # read (default LARGE) only for the WK_ makers (io.F); otherwise a
              # zero module variable, which leaves the source band empty
              Ywidth_WK=float(p.get("Ywidth_WK", LARGE)) if wk.startswith("WK_") else 0.0,
              Width_WK=0.0, D_gen=0.0, Beta_gen=0.0,
              rlamda=0.0, Tperiod=float(p.get("Tperiod", 0.0)), Time_ramp=float(p.get("Time_ramp", 0.0)),
              FreqPeak=0.0, NumFreq=0)
    if wk.startswith("WK_REG"):
        theta = float(p.get("Theta_WK", 0.0)) * PI_SP / 180.0
        if periodic and theta != 0.0:
            raise NotImplementedError("periodic WK_REG with an oblique angle")
        per = float(p["Tperiod"])
        d_gen, beta_gen, rl, wl = _wk_source_coeffs(1.0 / per, float(p["AMP_WK"]), theta,
                                                    float(p["DEP_WK"]), delta)
        st.update(D_gen=d_gen, Beta_gen=beta_gen, rlamda=rl, Width_WK=delta * wl / 2.0)
    elif wk.startswith("WK_DATA2D"):
        s = case.spectrum
        keep = np.abs(np.asarray(s["dire"])) < 60.0
        dire = np.asarray(s["dire"])[keep] * DEG2RAD
        amp = np.asarray(s["amp"])[:, keep]
        phase = np.asarray(s["phase_deg"])[:, keep] * 0.005555555555556 * PI_SP
        freq = np.asarray(s["freq"], np.float64)
        if periodic:
            raise NotImplementedError("periodic WK_DATA2D")
        h = float(p["DEP_WK"])
        F, D = np.meshgrid(freq, dire, indexing="ij")
        d_gen, beta_gen, rl, _ = _wk_source_coeffs(F, amp, D, h, delta)
        _, _, _, wl_peak = _wk_source_coeffs(1.0 / float(s["peak_period"]), 1.0, 0.0, h, delta)
        st.update(Width_WK=delta * wl_peak / 2.0, FreqPeak=1.0 / float(s["peak_period"]),
                  NumFreq=len(freq), OMGN2D=(2.0 * PI_SP * freq).astype(rdt))
        st["_wk2d"] = (d_gen, beta_gen[:, 0], rl, phase)
    elif not (wk.startswith("INI_SOL") or wk.startswith("NONE")):
        raise NotImplementedError(f"wavemaker {wk}")
    xmk = np.zeros(M)
    ymk = np.zeros(N)
    xmk[ib - 1:ie] = (np.arange(ib, ie + 1) - ib) * dx
    ymk[jb - 1:je] = (np.arange(jb, je + 1) - jb) * dy
    inx = np.where(np.abs(xmk[ib - 1:ie] - st["Xc_WK"]) < st["Width_WK"])[0]
    iny = np.where(np.abs(ymk[jb - 1:je] - st["Yc_WK"]) < st["Ywidth_WK"] / 2.0)[0]
    st["ilo"], st["ihi"] = (int(inx[0]) + ib, int(inx[-1]) + ib) if len(inx) else (ie + 1, ib - 1)
    st["jlo"], st["jhi"] = (int(iny[0]) + jb, int(iny[-1]) + jb) if len(iny) else (je + 1, jb - 1)
    st["xmk_wk"] = xmk.astype(rdt)
    st["ymk_wk"] = ymk.astype(rdt)
    if "_wk2d" in st:
        d_gen, beta1, rl, phase = st.pop("_wk2d")
        nf = st["NumFreq"]
        cm = np.zeros((nf, N, M), rdt)
        sm = np.zeros((nf, N, M), rdt)
        if st["ihi"] >= st["ilo"] and st["jhi"] >= st["jlo"]:
            I = np.arange(st["ilo"], st["ihi"] + 1)
            J = np.arange(st["jlo"], st["jhi"] + 1)
            xx = (I - ib) * dx - st["Xc_WK"]
            yy = (J - jb) * dy
            for kf in range(nf):  # CALCULATE_Cm_Sm, wavemaker.F:831-891
                ex = np.exp(-beta1[kf] * xx ** 2)[None, :]
                arg = rl[kf][:, None] * yy[None, :] + phase[kf][:, None]      # (ndir, nj)
                c = (d_gen[kf][:, None] * np.cos(arg)).sum(0)[:, None] * ex
                s_ = (d_gen[kf][:, None] * np.sin(arg)).sum(0)[:, None] * ex
                cm[kf, st["jlo"] - 1:st["jhi"], st["ilo"] - 1:st["ihi"]] = c
                sm[kf, st["jlo"] - 1:st["jhi"], st["ilo"] - 1:st["ihi"]] = s_
        st["Cm"], st["Sm"] = cm, sm

    # sponge (init.F:965-984 -> sponge.F:591)
    if st["DIRECT_SPONGE"]:
        st["SPONGE"] = _sponge(M, N, dx, dy, (float(p.get("Sponge_west_width", 0)),
                                              float(p.get("Sponge_east_width", 0)),
                                              float(p.get("Sponge_south_width", 0)),
                                              float(p.get("Sponge_north_width", 0))),
                               float(p.get("R_sponge", 0.85)), float(p.get("A_sponge", 5.0)))
    else:
        st["SPONGE"] = np.ones((N, M))
    widths = tuple(float(p.get(f"Sponge_{s}_width", 0)) for s in ("west", "east", "south", "north"))
    if st["FRICTION_SPONGE"]:  # init.F:980-989
        st["CD_4_SPONGE"] = _tanh_sponge(M, N, dx, dy, widths, float(p.get("CDsponge", 5.0))).astype(rdt)
    if st["DIFFUSION_SPONGE"]:  # init.F:973-978
        st["nu_sponge"] = _tanh_sponge(M, N, dx, dy, widths, float(p.get("Csp", 0.1))).astype(rdt)

    # initial surface and velocity: still water, or the solitary wave of
    # INITIAL_SOLITARY_WAVE (init.F:907-910, samples.F:511-571) over all local
    # cells, ghosts included, centred where I - 1 = XWAVEMAKER/DX
    eta = np.zeros((N, M))
    u0 = np.zeros((N, M))
    if wk.startswith("INI_SOL"):
        _, bue, ae1, ae2, au = _solitary_coeffs(float(p["AMP"]), float(p["DEP"]), rdt)
        I = np.arange(1, M + 1).astype(rdt)
        with np.errstate(over="ignore"):  # far from the crest cosh overflows: sech = 0
            sc = rdt(1) / np.cosh(bue * (I - rdt(float(p["XWAVEMAKER"])) / rdt(dx) - rdt(1)) * rdt(dx))
        e1 = ae1 * sc * sc + ae2 * sc * sc * sc * sc
        u1 = au * sc * sc
        if not bool(p.get("SolitaryPositiveDirection", True)):
            u1 = -u1
        eta[:, :] = e1.astype(np.float64)[None, :]
        u0[:, :] = u1.astype(np.float64)[None, :]
    # masks and depths at dry cells, init.F:1000-1025, 1143-1219
    mask = np.where(eta < -depth, 0, 1).astype(np.int32)
    eta = np.where(mask == 0, -mindep - depth, eta)
    h = np.maximum(eta * st["Gamma3"] + depth, mindep)
    # HU = H U, Ubar = HU + gamma1 U1p H (init.F:1017-1026); U1p is zero there
    # because CAL_DISPERSION runs before MASK9 is formed (:1158-1163)
    hu = h * u0
    mask9 = np.zeros((N, M), np.int32)
    m = mask
    prod = np.ones((N - 2, M - 2), np.int32)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            prod = prod * m[1 + dj:N - 1 + dj, 1 + di:M - 1 + di]
    mask9[jb - 1:je, ib - 1:ie] = prod[jb - 2:je - 1, ib - 2:ie - 1]
    dry = mask < 1
    # later cell in loop order wins (init.F:1210-1219 over 2..Mloc-1, 2..Nloc-1)
    inner = np.zeros_like(dry)
    inner[1:N - 1, 1:M - 1] = True
    d_in = dry & inner
    # face f lies between cells f-1 and f; cell f-1 writes Depth(f), cell f
    # writes Depth(f-1) afterwards (same row, later in loop order)
    fx = depx.copy()
    for f in range(1, M + 1):
        a = d_in[:, f - 1]
        fx[:, f] = np.where(a, depth[:, f] if f < M else fx[:, f], fx[:, f])
    for f in range(0, M):
        b = d_in[:, f]
        if f >= 1:
            fx[:, f] = np.where(b, depth[:, f - 1], fx[:, f])
    fy = depy.copy()
    for g in range(1, N + 1):
        a = d_in[g - 1, :]
        fy[g, :] = np.where(a, depth[g, :] if g < N else fy[g, :], fy[g, :])
    for g in range(0, N):
        b = d_in[g, :]
        if g >= 1:
            fy[g, :] = np.where(b, depth[g - 1, :], fy[g, :])
    st.update(Depth=depth.astype(rdt), DepthX=fx.astype(rdt), DepthY=fy.astype(rdt),
              Eta=eta.astype(rdt), H=h.astype(rdt), MASK=mask, MASK9=mask9,
              U=u0.astype(rdt), HU=hu.astype(rdt), Ubar=hu.astype(rdt),
              MASK_STRUC=np.ones((N, M), np.int32))
    for k in ("Cd", "SPONGE"):
        st[k] = st[k].astype(rdt)
    return st
