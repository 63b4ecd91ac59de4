"""GPU (libfwgpu.so through the C ABI) against the CPU oracle and the
reference's golden runs.

Per routine: starting from the same state, every routine of three whole
steps is run on the GPU and in the C restatement; every array the step
touches must agree bit for bit.  Whole runs: fw_step over hundreds of steps
must reproduce the compiled reference's state bit for bit (fixtures)."""
import numpy as np
import pytest

from common import STATE_KEYS, bitwise_mismatch, load_golden, oracle_available, rel_l2

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle_available(), reason="oracle not built")]

CASES = ["tiny2d", "tiny1d", "island", "tiny2d_dx15", "config0", "member3", "tiny2d_fsponge", "tiny2d_nobrk",
         "tiny2d_swe", "tiny2d_dy1", "tiny2d_dy15"]


def _solvers(name, prec, full_state=True):
    import fwgpu
    import oracle
    init, steps = load_golden(name, prec)
    return fwgpu.Solver(init, prec, full_state=full_state), oracle.OracleSolver(init), init, steps


# arrays the production path (full_state off) does not keep up to date: face
# states and wave speeds at interior faces, consumed inside FLUXES only
FACE_STATES = {"UxL", "UxR", "VxL", "VxR", "HUxL", "HUxR", "HVxL", "HVxR", "EtaRxL", "EtaRxR",
               "HxL", "HxR", "U4xL", "U4xR", "V4xL", "V4xR", "PL", "PR", "FxL", "FxR", "GxL",
               "GxR", "SxL", "SxR", "UyL", "UyR", "VyL", "VyR", "HUyL", "HUyR", "HVyL", "HVyR",
               "EtaRyL", "EtaRyR", "HyL", "HyR", "U4yL", "U4yR", "V4yL", "V4yR", "QL", "QR",
               "FyL", "FyR", "GyL", "GyR", "SyL", "SyR"}
# and the first time/space derivatives of CAL_DISPERSION, consumed only at
# their own cell by the nonlinear dispersive terms
DISP_FIRST = {"Ut", "Vt", "Ux", "Vx", "Uy", "Vy", "DUx", "DVx", "DUy", "DVy", "Utx", "Vty",
              "Utxx", "Utxy", "Vtxy", "Vtyy", "DUtx", "DVty", "DUtxx", "DUtxy", "DVtxy", "DVtyy"}
# and the source terms, formed inside the RK update of their own cell
SOURCES = {"SourceX", "SourceY", "WaveMaker_Mass", "BreakSourceX", "BreakSourceY", "FrcInsX", "FrcInsY"}
# and the mean-flow diagnostics of ESTIMATE_HUV and the roller flux and
# undertow of WAVE_BREAKING, read by no routine of the step
DIAGNOSTICS = {"P_center", "Q_center", "U_davg", "V_davg", "Wsurf", "ROLLER_FLUX", "UNDERTOW_U", "UNDERTOW_V"}
# and the nonlinear dispersive terms that SourceTerms reads only as the sums
# U1pp+U2+U3, V1pp+V2+V3 (kept as such)
DISP_SUMMED = {"U1pp", "V1pp", "U2", "V2", "U3", "V3"}
# and the second derivatives read only by CAL_DISPERSION's own terms (the
# one-pass dispersion kernel keeps them on chip; the cross derivatives,
# which ESTIMATE_HUV reads, are stored)
DISP_SECOND = {"Uxx", "DUxx", "Vyy", "DVyy"}
PRODUCTION_SKIP = FACE_STATES | DISP_FIRST | SOURCES | DIAGNOSTICS | DISP_SUMMED | DISP_SECOND
# and, in fw_step's production path, the cross derivatives and U1p, V1p,
# which the RK update forms itself at its own cell (the one-pass dispersion
# does not store them; the per-routine CAL_DISPERSION does)
DISP_CROSS = {"Uxy", "Vxy", "DUxy", "DVxy", "U1p", "V1p"}


def _compare(g, o, where, full_state=True, step=False):
    import fwgpu
    bad = {}
    for k in fwgpu.ALL_FIELDS:
        if k in fwgpu.AUX or (not full_state and k in PRODUCTION_SKIP):
            continue
        if step and not full_state and k in DISP_CROSS:
            continue
        n = bitwise_mismatch(g.download(k), o[k])
        if n:
            bad[k] = n
    assert not bad, f"{where}: {bad}"


@pytest.mark.parametrize("full_state", [True, False])
@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("name", CASES)
def test_each_routine_matches_oracle(name, prec, full_state, gpu_available):
    print(f"\n[routine parity] case={name} prec={prec} full_state={full_state}", flush=True)
    g, o, init, _ = _solvers(name, prec, full_state)
    cmp = lambda where: _compare(g, o, where, full_state)  # noqa: E731
    cmp("upload")
    _routine_steps(g, o, init, cmp, 3)
    g.close()


def _routine_steps(g, o, init, cmp, nsteps):
    """nsteps whole steps routine by routine (the order of fw_step), every
    array compared after each routine."""
    for step in range(1, nsteps + 1):
        g.save_stage0(1)
        for a, b in (("Eta0", "Eta"), ("Ubar0", "Ubar"), ("Vbar0", "Vbar")):
            o[a][...] = o[b]
        g.exchange(); o.exchange(); cmp(f"s{step} exchange")
        gdt, gt = g.estimate_dt(); o.estimate_dt()
        assert (gdt, gt) == (o.dt.value, o.time.value)
        g.save_stage0(2)
        o["U0"][...] = o["U"]; o["V0"][...] = o["V"]
        for st in (1, 2, 3):
            if init["DISPERSION"]:
                g.dispersion(); o.dispersion(); cmp(f"s{step}.{st} dispersion")
            g.fluxes(); o.fluxes(); cmp(f"s{step}.{st} fluxes")
            g.sources(); o.sources(); cmp(f"s{step}.{st} sources")
            g.estimate_huv(st); o.estimate_huv(st); cmp(f"s{step}.{st} huv")
            g.update_mask(); o.update_mask(); cmp(f"s{step}.{st} mask")
            g.breaking(); o.breaking(); cmp(f"s{step}.{st} breaking")
            g.exchange(); o.exchange(); cmp(f"s{step}.{st} exchange")
            if init["DIRECT_SPONGE"]:
                g.sponge(); o.sponge(); cmp(f"s{step}.{st} sponge")


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_periodic_2d_matches_oracle(prec, gpu_available):
    """A periodic domain 40 rows tall with y variation (shoal off the centre
    line), built from the case description: 100 fused steps (every array bit
    for bit after steps 1 and 100), then two steps routine by routine."""
    import cases
    import fwgpu
    import fwsetup
    import oracle
    print(f"\n[periodic 2-D parity] 120 x 40 {prec}", flush=True)
    st = fwsetup.build_state(cases.periodic2d(), prec)
    assert st["PERIODIC"]
    g = fwgpu.Solver(st, prec, full_state=True)
    o = oracle.OracleSolver(st)
    for n in (1, 99):
        gdt, gt = g.step(n)
        o.run(n)
        assert (gdt, gt) == (o.dt.value, o.time.value)
        _compare(g, o, f"after {n} more fused steps")
    v = o["V"]
    assert np.abs(v).max() > 1e-3 and np.abs(np.diff(v, axis=0)).max() > 1e-4  # V varies along y
    assert int((o["MASK"] == 0).sum()) > 0
    _routine_steps(g, o, st, lambda where: _compare(g, o, where), 2)
    g.close()


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_ragged_grid_matches_oracle(prec, gpu_available):
    """75 x 53 cells, built from the case description (no fixture): partial
    tiles, partial blocks of lines, partial 64-row chunks and rows past the
    last 8-row group in both directions of the line solves.  40 fused steps
    (every array bit for bit after steps 1 and 40), then two steps routine
    by routine."""
    import cases
    import fwgpu
    import fwsetup
    import oracle
    print(f"\n[ragged-grid parity] 75 x 53 {prec}", flush=True)
    st = fwsetup.build_state(cases.ragged2d(), prec)
    g = fwgpu.Solver(st, prec, full_state=True)
    o = oracle.OracleSolver(st)
    for n in (1, 39):
        gdt, gt = g.step(n)
        o.run(n)
        assert (gdt, gt) == (o.dt.value, o.time.value)
        _compare(g, o, f"after {n} more fused steps")
    assert int((o["MASK"] == 0).sum()) > 0  # the wet/dry front is on the grid
    _routine_steps(g, o, st, lambda where: _compare(g, o, where), 2)
    g.close()


@pytest.mark.parametrize("full_state", [True, False])
@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("name", CASES)
def test_fused_step_matches_oracle(name, prec, full_state, gpu_available):
    """fw_step (the whole-step path, with its fused kernels) against the
    oracle's whole steps: every array bit for bit after 1 and 3 steps."""
    print(f"\n[fused-step parity] case={name} prec={prec} full_state={full_state}", flush=True)
    g, o, init, _ = _solvers(name, prec, full_state)
    for n in (1, 2):
        gdt, gt = g.step(n)
        o.run(n)
        assert (gdt, gt) == (o.dt.value, o.time.value)
        _compare(g, o, f"after fw_step({n})", full_state, step=True)
    g.close()


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("name", ["tiny2d", "island"])
def test_fluxes_with_signed_zeros(name, prec, gpu_available):
    """FLUXES on a state where 40 % of the cells of every reconstructed field
    hold +0 or -0 (a masked negative velocity is -0): slopes such as
    (-0) - (+0) = -0 beside nonzero ones, where only MIN/MAX tie-breaking
    decides the sign of a zero limited slope.  Every array bit for bit."""
    import fwgpu
    import oracle
    print(f"\n[signed zeros] case={name} prec={prec}", flush=True)
    init, steps = load_golden(name, prec)
    st = dict(init)
    st.update({k: v for k, v in steps[sorted(steps)[1]].items() if not np.isscalar(v)})
    rng = np.random.default_rng(7)
    for k in ("U", "V", "HU", "HV", "Eta"):
        a = np.array(st[k], copy=True)
        r = rng.random(a.shape)
        a[r < 0.2] = 0.0
        a[(r >= 0.2) & (r < 0.4)] = -0.0
        st[k] = a
    g, o = fwgpu.Solver(st, prec, full_state=True), oracle.OracleSolver(st)
    _compare(g, o, "upload")
    g.exchange(); o.exchange()
    gdt, gt = g.estimate_dt(); o.estimate_dt()
    assert (gdt, gt) == (o.dt.value, o.time.value)
    g.save_stage0(2)
    o["U0"][...] = o["U"]; o["V0"][...] = o["V"]
    if init["DISPERSION"]:
        g.dispersion(); o.dispersion()
    g.fluxes(); o.fluxes()
    _compare(g, o, "fluxes")
    # the case is exercised: some slope (-0) - (+0) sits between wet cells
    u, m = st["U"], st["MASK"]
    t2 = u[:, 1:] - u[:, :-1]
    assert np.any(np.signbit(t2) & (t2 == 0) & (m[:, 1:] == 1) & (m[:, :-1] == 1))
    g.close()


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_fluxes_with_slopes_beyond_quotient_range(prec, gpu_available):
    """FLUXES where 2 % of the cells of HU and HV hold +-1e10, so that some
    limited slopes exceed 2^30: the FP32 van Leer weights then take the
    correctly rounded division instead of the remainder-corrected quotient.
    Every array bit for bit against the oracle."""
    import fwgpu
    import oracle
    print(f"\n[large slopes] prec={prec}", flush=True)
    init, steps = load_golden("tiny2d", prec)
    st = dict(init)
    st.update({k: v for k, v in steps[sorted(steps)[1]].items() if not np.isscalar(v)})
    rng = np.random.default_rng(5)
    for k in ("HU", "HV"):
        a = np.array(st[k], copy=True)
        sel = rng.random(a.shape) < 0.02
        a[sel] = np.where(rng.random(a.shape) < 0.5, 1e10, -1e10)[sel].astype(a.dtype)
        st[k] = a
    g, o = fwgpu.Solver(st, prec, full_state=True), oracle.OracleSolver(st)
    g.exchange(); o.exchange()
    gdt, gt = g.estimate_dt(); o.estimate_dt()
    assert (gdt, gt) == (o.dt.value, o.time.value)
    g.save_stage0(2)
    o["U0"][...] = o["U"]; o["V0"][...] = o["V"]
    g.fluxes(); o.fluxes()
    _compare(g, o, "fluxes")
    assert np.nanmax(np.abs(o["HUxL"])) > 2.0 ** 30  # the large slopes reach the states
    g.close()


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_full_size_beach_step_matches_oracle(prec, gpu_available):
    """The benchmark workload itself (2048 x 2048 beach, WK_DATA2D waves,
    sponges, dispersion and breaking; production mode): one fw_step against
    one oracle step, every array the production path keeps, bit for bit
    (the oracle takes 5-15 s for the step)."""
    import fwcases
    import fwgpu
    import fwsetup
    import oracle
    print(f"\n[full-size parity] beach 2048 x 2048 {prec}, 1 step", flush=True)
    st = fwsetup.build_state(fwcases.beach_2d(mglob=2048, nglob=2048, dx=2.0), prec)
    g = fwgpu.Solver(st, prec)
    gdt, gt = g.step(1)
    g.synchronize()
    o = oracle.OracleSolver(st)
    o.run(1)
    assert (gdt, gt) == (o.dt.value, o.time.value)
    _compare(g, o, "after fw_step(1) at 2048^2", full_state=False, step=True)
    g.close()


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("name", ["tiny2d", "ragged2d"])
def test_fused_routine_and_upload_interleaved(name, prec, gpu_available):
    """Production fw_step mixed with the per-routine entry points and an
    upload: fused steps while the wet/dry front moves, a step routine by
    routine, fused steps, an upload of a changed MASK9 (which the line
    coefficients depend on), fused steps: every array the production path
    keeps, bit for bit against the oracle."""
    import cases
    import fwgpu
    import fwsetup
    import oracle
    print(f"\n[fused, routine and upload interleaved] case={name} {prec}", flush=True)
    if name == "ragged2d":
        st = fwsetup.build_state(cases.ragged2d(), prec)
    else:
        st, _ = load_golden(name, prec)
    g = fwgpu.Solver(st, prec)
    o = oracle.OracleSolver(st)
    cmp = lambda where: _compare(g, o, where, full_state=False, step=True)  # noqa: E731
    masks = [o["MASK"].copy()]
    for n in (1, 4, 15):
        gdt, gt = g.step(n)
        o.run(n)
        assert (gdt, gt) == (o.dt.value, o.time.value)
        cmp(f"after fw_step({n})")
        masks.append(o["MASK"].copy())
    assert any((a != b).any() for a, b in zip(masks, masks[1:])), "no wet/dry change in the fused steps"
    _routine_steps(g, o, st, cmp, 1)
    for n in (1, 2):
        gdt, gt = g.step(n)
        o.run(n)
        assert (gdt, gt) == (o.dt.value, o.time.value)
        cmp(f"after routines, fw_step({n})")
    m9 = o["MASK9"].copy()
    wet = np.argwhere(m9 == 1)
    sel = wet[:: max(1, len(wet) // 40)]
    m9[sel[:, 0], sel[:, 1]] = 0
    o["MASK9"][...] = m9
    g.upload("MASK9", m9)
    for n in (1, 2):
        gdt, gt = g.step(n)
        o.run(n)
        assert (gdt, gt) == (o.dt.value, o.time.value)
        cmp(f"after a MASK9 upload, fw_step({n})")
    g.close()


TOL = {"sp": 1e-5, "dp": 1e-10}


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("name", CASES)
def test_whole_run_matches_reference(name, prec, gpu_available):
    """fw_step over the fixture's steps against the compiled reference (every
    fixture reaches step 1000; config0 also its whole 60 s run, 1711 steps),
    with the Froude cap's ATAN2/COS/SIN evaluated by the host's C library
    (FW_OPT_HOST_MATH) as in the reference: every state array bit for bit at
    every checked step, which meets the bar of identical wet/dry masks and
    eta/u/v relative L2 within 1e-5 (FP32) / 1e-10 (FP64) at step 1000."""
    import fwgpu
    print(f"\n[whole-run parity] case={name} prec={prec}", flush=True)
    init, steps = load_golden(name, prec)
    g = fwgpu.Solver(init, prec, host_math=True)
    assert max(steps) >= 1000
    done = 0
    for s in sorted(steps):
        dt, t = g.step(s - done)
        done = s
        ref = steps[s]
        assert (dt, t) == (ref["DT"], ref["TIME"]), (s, dt, ref["DT"])
        nbit = {k: bitwise_mismatch(g.download(k), ref[k]) for k in STATE_KEYS}
        assert not any(nbit.values()), (s, {k: v for k, v in nbit.items() if v})
        print(f"  step {s}: every state array bit for bit")
    g.close()


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("name", CASES)
def test_whole_run_device_math(name, prec, gpu_available):
    """The production path (device math library for the Froude cap) over
    the same runs: bitwise the reference until the cap first acts; from then
    on last-bit differences of ATAN2/COS/SIN can grow in the strongly forced
    breaking runs.  Masks identical and eta/u/v relative L2 within 1e-5 /
    1e-10 are required up to step 250; the differences at the later fixture
    steps are printed (their bound is the FP32 test below: no larger than
    the reference's own FP32 error against its FP64 build)."""
    import fwgpu
    print(f"\n[whole-run, device math] case={name} prec={prec}", flush=True)
    init, steps = load_golden(name, prec)
    g = fwgpu.Solver(init, prec)
    done = 0
    for s in sorted(steps):
        dt, t = g.step(s - done)
        done = s
        ref = steps[s]
        masks = bitwise_mismatch(g.download("MASK"), ref["MASK"])
        errs = {k: rel_l2(g.download(k)[3:-3, 3:-3], ref[k][3:-3, 3:-3]) for k in ("Eta", "U", "V")}
        nbit = {k: bitwise_mismatch(g.download(k), ref[k]) for k in STATE_KEYS}
        print(f"  step {s}: mask cells differing {masks}; rel L2 {errs}; arrays with non-identical bits: "
              f"{dict((k, v) for k, v in nbit.items() if v)}")
        if s <= 250:
            assert masks == 0, f"mask differs at step {s}"
            assert all(e <= TOL[prec] for e in errs.values()), (s, errs)
    g.close()


@pytest.mark.parametrize("name", CASES)
def test_fp32_error_against_fp64_reference(name, gpu_available):
    """The FP32 GPU state against the FP64 reference at the last fixture step
    (the error stated for every FP32 result).  With the Froude cap on the
    host's C library the GPU's FP32 run is the reference's FP32 run bit for
    bit, so its error is exactly the reference's own; the production path
    (device math library) is printed beside it: in the strongly forced
    breaking runs its last-bit differences grow like FP32 rounding does, so
    after 1000 steps its error is of the size of the reference's, not equal
    to it."""
    import fwgpu
    init32, steps32 = load_golden(name, "sp")
    _, steps64 = load_golden(name, "dp")
    last = max(steps32)
    ref = steps64[last]
    inner = lambda a: a[3:-3, 3:-3]  # noqa: E731
    e_ref = rel_l2(inner(steps32[last]["Eta"]), inner(ref["Eta"]))
    errs = {}
    for host_math in (True, False):
        g = fwgpu.Solver(init32, "sp", host_math=host_math)
        g.step(last)
        errs[host_math] = (rel_l2(inner(g.download("Eta")), inner(ref["Eta"])),
                           bitwise_mismatch(g.download("MASK"), ref["MASK"]))
        g.close()
    print(f"\n[fp32 vs fp64 reference] {name} after {last} steps: eta rel L2 reference-FP32 = {e_ref:.3e}, "
          f"GPU-FP32 host math = {errs[True][0]:.3e}, device math = {errs[False][0]:.3e}; mask cells "
          f"differing from FP64: host math {errs[True][1]}, device math {errs[False][1]}")
    assert errs[True][0] == e_ref


def test_blowup_is_reported(gpu_available):
    import fwgpu
    init, _ = load_golden("tiny2d", "sp")
    init = dict(init)
    init["EtaBlowVal"] = 1e-9
    g = fwgpu.Solver(init, "sp")
    with pytest.raises(fwgpu.FwError) as e:
        g.step(3)
    assert e.value.code == fwgpu.FW_E_BLOWUP
    g.close()


def test_fp64_budget_matches_host_sum(gpu_available):
    import fwgpu
    init, _ = load_golden("tiny2d", "sp")
    g = fwgpu.Solver(init, "sp")
    g.step(40)
    s = g.stats()
    eta = g.download("Eta")[3:-3, 3:-3].astype(np.float64)
    mask = g.download("MASK")[3:-3, 3:-3]
    dxdy = float(init["DX"]) * float(init["DY"])
    wet = mask > 0
    mass = (eta * wet).sum() * dxdy
    assert s["wet_cells"] == int(wet.sum())
    assert s["mass_volume"] == pytest.approx(mass, rel=1e-12, abs=1e-9)
    # energy: 0.5 (eta + h) (u^2 + v^2) + 0.5 g eta^2 over wet cells, total
    # depth formed in FP64 from the FP32 state
    u = g.download("U")[3:-3, 3:-3].astype(np.float64)
    v = g.download("V")[3:-3, 3:-3].astype(np.float64)
    h = eta + g.download("Depth")[3:-3, 3:-3].astype(np.float64)
    en = ((0.5 * h * (u * u + v * v) + 0.5 * 9.81 * eta * eta) * wet).sum() * dxdy
    assert s["energy"] == pytest.approx(en, rel=1e-12)
    assert s["max_eta"] == eta[wet].max() and s["min_eta"] == eta[wet].min()
    assert s["max_u"] == np.abs(u[wet]).max() and s["max_v"] == np.abs(v[wet]).max()
    g.close()
