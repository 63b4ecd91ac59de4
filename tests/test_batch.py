"""Ensembles (fw_batch_*): members of uneven size advanced by one launch per
phase of the step, replayed from captured graphs, against the same members
advanced alone with fw_step.  The batched path runs the per-domain kernels'
own bodies with each member's own tile of the launch, so every array the
production path keeps must agree bit for bit; a member that blows up or
whose fixed time step would halve stops alone."""
import numpy as np
import pytest

from common import bitwise_mismatch

pytestmark = pytest.mark.gpu

KEYS = ["Eta", "Eta0", "U", "V", "U0", "V0", "HU", "HV", "Ubar", "Vbar", "Ubar0", "Vbar0", "H", "MASK", "MASK9",
        "AGE_BREAKING", "nu_break", "ETAT", "ETAx", "ETAy", "U4", "V4", "DepthX", "DepthY", "P", "Q",
        "Fx", "Fy", "Gx", "Gy"]


def fixed_dt_state(case, prec, frac=0.5):
    """Host state with FIXED_DT: half the CFL step of the initial state,
    rounded down to 1 ms (kept all run long by these short runs)."""
    import fwgpu
    import fwsetup
    st = fwsetup.build_state(case, prec)
    g = fwgpu.Solver(st, prec)
    dt, _ = g.estimate_dt()
    g.close()
    dtf = float(np.floor(1000.0 * frac * dt) / 1000.0)
    case.params["DT_fixed"] = dtf
    st = fwsetup.build_state(case, prec)
    assert st["FIXED_DT"] == 1 and float(st["DT_fixed"]) == dtf
    return st


def members(prec, kinds=("reg", "data2d", "sol"), n_each=2, seed0=0):
    """Small members of every wavemaker kind: WK_REG beaches with bumps,
    WK_DATA2D spectral beaches, INI_SOL islands; dx 2 or 4 m."""
    import fwcases
    out = []
    for k in range(n_each):
        s = seed0 + k
        if "reg" in kinds:
            out.append(fixed_dt_state(fwcases.member(s, mrange=(40, 130), nrange=(24, 90)), prec))
        if "data2d" in kinds:
            c = fwcases.beach_2d(mglob=48 + 17 * k, nglob=40 + 9 * k, dx=2.0, nfreq=4, ndir=3, flat_depth=3.0,
                                 slope=1 / 25.0, x_toe=40.0, xc_wk=20.0, sponge=10.0)
            out.append(fixed_dt_state(c, prec))
        if "sol" in kinds:
            out.append(fixed_dt_state(fwcases.island(s, dx=4.0, mglob=50 + 13 * k, nglob=36 + 7 * k,
                                                     r_base=50.0, h0=3.0), prec))
    return out


def run_alone(states, prec, nsteps):
    import fwgpu
    res = []
    for st in states:
        g = fwgpu.Solver(st, prec)
        g.step(nsteps)
        res.append(({k: g.download(k) for k in KEYS}, g.clock))
        g.close()
    return res


def compare(batch_members, alone, where):
    for m, (g, (ref, clk)) in enumerate(zip(batch_members, alone)):
        assert g.clock == clk, (where, m, g.clock, clk)
        bad = {k: bitwise_mismatch(g.download(k), v) for k, v in ref.items()}
        bad = {k: v for k, v in bad.items() if v}
        assert not bad, f"{where}: member {m}: {bad}"


@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_batch_matches_members_alone(prec, gpu_available):
    """Six members of three wavemaker kinds and six sizes (40..130 x 24..90
    cells): 23 batched steps (graphs, an odd count: the buffer parity ends
    swapped) bit for bit equal to 23 steps of each member alone."""
    import fwgpu
    print(f"\n[batch parity] {prec}", flush=True)
    sts = members(prec)
    alone = run_alone(sts, prec, 23)
    gs = [fwgpu.Solver(st, prec) for st in sts]
    b = fwgpu.Batch(gs)
    b.step(10)
    b.step(13)
    assert b.status() == [0] * len(gs)
    compare(gs, alone, "after 23 batched steps")
    b.close()
    for g in gs:
        g.close()


def test_batch_eager_equals_graph_and_continues_alone(gpu_available):
    """Launching the phases eagerly gives the same bits as replaying the
    captured graphs; a member continued with fw_step after the batch (and
    the batch continued after per-member steps) stays on the same path."""
    import fwgpu
    sts = members("sp", kinds=("reg", "sol"), n_each=2)
    alone = run_alone(sts, "sp", 9)
    gs = [fwgpu.Solver(st, "sp") for st in sts]
    b = fwgpu.Batch(gs)
    b.set_graph(False)
    b.step(3)
    for g in gs:
        g.step(3)              # members alone in between
    b.set_graph(True)
    b.step(3)
    compare(gs, alone, "eager + alone + graph")
    b.close()
    for g in gs:
        g.close()


def test_batch_periodic_members(gpu_available):
    """Periodic quasi-1-D beaches (Sherman-Morrison line solves per thread,
    periodic ghost fills) of three lengths."""
    import fwcases
    import fwgpu
    sts = [fixed_dt_state(fwcases.beach_1d(mglob=m, nglob=4, xslp=0.4 * m, xc_wk=0.2 * m, sponge_west=0.1 * m),
                          "sp") for m in (150, 200, 260)]
    alone = run_alone(sts, "sp", 15)
    gs = [fwgpu.Solver(st, "sp") for st in sts]
    b = fwgpu.Batch(gs)
    b.step(15)
    compare(gs, alone, "periodic batch")
    b.close()
    for g in gs:
        g.close()


def test_member_blowup_and_dt_change_stop_alone(gpu_available):
    """Member 0 blows up at its first step (EtaBlowVal tiny): it stops after
    that step.  Member 1's fixed step is far above its CFL limit, so the
    first ESTIMATE_DT would halve it: it stops before any step, its state
    untouched.  The others run all 8 steps as if alone."""
    import fwgpu
    sts = members("sp", kinds=("reg",), n_each=4, seed0=10)
    sts[0] = dict(sts[0], EtaBlowVal=1e-9)
    sts[1] = dict(sts[1], DT_fixed=50.0 * float(sts[1]["DT_fixed"]))
    alone = run_alone(sts[2:], "sp", 8)
    gs = [fwgpu.Solver(st, "sp") for st in sts]
    eta1 = gs[1].download("Eta")
    b = fwgpu.Batch(gs)
    with pytest.raises(fwgpu.FwError):
        b.step(8)
    assert b.status()[:2] == [fwgpu.FW_BATCH_BLOWUP, fwgpu.FW_BATCH_DT_CHANGED]
    assert b.status()[2:] == [0, 0]
    compare(gs[2:], alone, "members that kept running")
    # member 0 took exactly one step: one step alone, blow-up then
    g0 = fwgpu.Solver(sts[0], "sp")
    with pytest.raises(fwgpu.FwError):
        g0.step(1)
    assert gs[0].clock == g0.clock
    assert bitwise_mismatch(gs[0].download("Eta"), g0.download("Eta")) == 0
    assert gs[1].clock[1] == 0.0
    assert bitwise_mismatch(gs[1].download("Eta")[3:-3, 3:-3], eta1[3:-3, 3:-3]) == 0
    b.close()
    for g in gs + [g0]:
        g.close()


def test_autotune_leaves_state_untouched(gpu_available):
    """The autotuner times trial steps of every candidate shape on a
    snapshot and restores it: afterwards the batch still produces the
    members' own results bit for bit, with the shapes it chose."""
    import fwgpu
    sts = members("sp", kinds=("reg", "data2d"), n_each=2, seed0=20)
    alone = run_alone(sts, "sp", 6)
    gs = [fwgpu.Solver(st, "sp") for st in sts]
    b = fwgpu.Batch(gs)
    rep = b.autotune(trial_steps=2)
    print("\n" + rep)
    assert "->" in rep
    assert all(g.clock[1] == 0.0 for g in gs)
    b.step(6)
    compare(gs, alone, "after autotune")
    print(b.shapes())
    b.close()
    for g in gs:
        g.close()
