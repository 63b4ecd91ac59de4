"""GPU (libfwgpu.so through the C ABI) against the oracle on the benchmark
workloads at their own sizes, and CHECK_BLOWUP's maximum.

* the 2048 x 2048 beach (the Mcell-steps/s workload): 10 whole steps, every
  array the production path keeps, bit for bit, FP32 and FP64;
* an ensemble member of the batched-ensemble distribution (member seed
  1000, 412 x 462 cells, dx = 2 m, WK_REG on a 2-D beach with bumps) and a
  conical-island member of the learned-ensemble distribution at the coarse
  spacing dx = 4 m (member seed 9000, 597 x 221 cells): whole steps bit for
  bit;
* the learned correction's network on a 1024 x 1024 coarse grid against the
  numpy checker on windows (the network is local: four 3x3 layers see four
  cells around each output cell);
* CHECK_BLOWUP (misc.F:394-433): fw_check_blowup's max |eta| equals the
  oracle's interior maximum bit for bit.
"""
import numpy as np
import pytest

from common import bitwise_mismatch, load_golden, oracle_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle_available(), reason="oracle not built")]

# arrays the production path keeps up to date after whole steps
PRODUCTION_KEYS = ["Eta", "Eta0", "U", "V", "U0", "V0", "HU", "HV", "Ubar", "Vbar", "Ubar0", "Vbar0", "H",
                   "MASK", "MASK9", "AGE_BREAKING", "nu_break", "ETAT", "ETAx", "ETAy", "U4", "V4",
                   "DepthX", "DepthY", "P", "Q", "Fx", "Fy", "Gx", "Gy"]


def _compare(g, o, keys, where):
    bad = {}
    for k in keys:
        n = bitwise_mismatch(g.download(k), o[k])
        if n:
            bad[k] = n
    assert not bad, f"{where}: {bad}"


def _run_compare(st, prec, chunks, where):
    import fwgpu
    import oracle
    g = fwgpu.Solver(st, prec)
    o = oracle.OracleSolver(st)
    done = 0
    for n in chunks:
        gdt, gt = g.step(n)
        assert o.run(n) == 0
        done += n
        assert (gdt, gt) == (o.dt.value, o.time.value), (done, gdt, o.dt.value)
        _compare(g, o, PRODUCTION_KEYS, f"{where} after {done} steps")
        print(f"  {where}: {done} steps bit for bit (dt={gdt:.6g})", flush=True)
    g.close()
    return o


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_full_size_beach_ten_steps(prec, gpu_available):
    """The 2048 x 2048 beach (WK_DATA2D 200-component waves, three sponges,
    dispersion, breaking, wetting-drying), 10 fused steps against 10 oracle
    steps, compared after steps 1, 5 and 10."""
    import fwcases
    import fwsetup
    print(f"\n[full-size parity] beach 2048 x 2048 {prec}, 10 steps", flush=True)
    st = fwsetup.build_state(fwcases.beach_2d(mglob=2048, nglob=2048, dx=2.0), prec)
    o = _run_compare(st, prec, (1, 4, 5), "beach 2048^2")
    assert int((o["MASK"] == 0).sum()) > 0


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_ensemble_member_full_size(prec, gpu_available):
    """Member seed 1000 of the batched-ensemble distribution at its own size
    (412 x 462 cells, dx = 2 m): 60 fused steps bit for bit."""
    import fwcases
    import fwsetup
    c = fwcases.member(0)
    print(f"\n[ensemble member] {c.params['Mglob']} x {c.params['Nglob']} {prec}", flush=True)
    st = fwsetup.build_state(c, prec)
    _run_compare(st, prec, (1, 29, 30), "member 1000")


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_island_member_coarse_spacing(prec, gpu_available):
    """Member seed 9000 of the learned-ensemble distribution (solitary wave
    over a conical island, INI_SOL) on the coarse grid, dx = 4 m, at its own
    size (597 x 221): 150 fused steps bit for bit."""
    import fwcases
    import fwsetup
    c = fwcases.island(0, dx=4.0)
    print(f"\n[island member] {c.params['Mglob']} x {c.params['Nglob']} dx=4 {prec}", flush=True)
    st = fwsetup.build_state(c, prec)
    _run_compare(st, prec, (1, 49, 100), "island 9000")


CASES = ["tiny2d", "tiny1d", "island", "config0", "member3", "tiny2d_nobrk"]


@pytest.mark.parametrize("prec", ["sp", "dp"])
@pytest.mark.parametrize("name", CASES + ["ragged2d"])
def test_check_blowup_maximum_matches_oracle(name, prec, gpu_available):
    """fw_check_blowup's max |eta| over the interior (misc.F:394-433) equals
    the oracle's after 0, 7 and 40 steps (the ragged 75 x 53 grid has partial
    rows and tiles)."""
    import cases
    import fwgpu
    import fwsetup
    import oracle
    if name == "ragged2d":
        st = fwsetup.build_state(cases.ragged2d(), prec)
    else:
        st, _ = load_golden(name, prec)
    g = fwgpu.Solver(st, prec)
    o = oracle.OracleSolver(st)
    for n in (0, 7, 33):
        if n:
            g.step(n)
            o.run(n)
        rc, m = g.check_blowup()
        assert rc == 0
        ref = o.lib.fwo_max_abs_eta(o.p, o.f)
        interior = float(np.abs(o["Eta"][3:-3, 3:-3]).max())
        assert m == ref == interior, (name, n, m, ref, interior)
    g.close()


@pytest.mark.timeout(300)
def test_network_on_1024_grid_matches_numpy(gpu_available):
    """The correction network over a whole 1024 x 1024 coarse grid (the
    learned workload's size), checked against the numpy restatement on nine
    64 x 64 windows: corners, edges and interior.  Each window is evaluated
    in numpy from the input window grown by four cells (the network's reach),
    clipped at the grid's edge, where both pad with zeros."""
    import cnn_numpy as cn
    import fwcnn
    from test_cnn import random_params, rel_l2
    n, w, r = 1024, 64, 4
    rng = np.random.default_rng(11)
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = np.stack([np.sin(0.01 * xx + 0.02 * yy), np.cos(0.015 * yy), 0.3 * np.sin(0.03 * xx),
                  1.0 + 0.001 * xx, (rng.random((n, n)) > 0.1).astype(np.float64)])[None].astype(np.float32)
    x += 0.05 * rng.standard_normal(x.shape).astype(np.float32)
    p = random_params(12)
    print("\n[cnn 1024^2] forward over the whole grid", flush=True)
    net = fwcnn.Net(n, n, 1, fwcnn.flatten(p))
    out = net.forward(x)
    net.close()
    for y0 in (0, 480, n - w):
        for x0 in (0, 512, n - w):
            ys, ye = max(0, y0 - r), min(n, y0 + w + r)
            xs, xe = max(0, x0 - r), min(n, x0 + w + r)
            ref, _ = cn.forward(p, x[:, :, ys:ye, xs:xe])
            ref = ref[:, :, y0 - ys:y0 - ys + w, x0 - xs:x0 - xs + w]
            e = rel_l2(out[:, :, y0:y0 + w, x0:x0 + w], ref)
            print(f"  window ({y0},{x0}): rel L2 {e:.2e}")
            assert e < 1e-3, (y0, x0, e)


def test_break_iteration_error_is_cleared(gpu_available):
    """WAVE_BREAKING reports FW_E_BREAK_ITER when the in-order age
    propagation does not settle, which NaN ages force (a candidate whose own
    age is NaN never compares equal to itself, so every sweep counts as a
    change until the sweep limit).  The error belongs to the step that raised
    it: after clean ages are uploaded again the next step succeeds."""
    import fwgpu
    init, _ = load_golden("tiny2d", "sp")
    g = fwgpu.Solver(init, "sp")
    g.step(250)                      # breaking is active on the beach by now
    age = g.download("AGE_BREAKING")
    g.upload("AGE_BREAKING", np.full_like(age, np.nan))
    with pytest.raises(fwgpu.FwError) as e:
        g.step(1)
    assert e.value.code == fwgpu.FW_E_BREAK_ITER
    g.upload("AGE_BREAKING", age)
    g.step(2)                        # no stale error from the flag
    assert np.isfinite(g.download("AGE_BREAKING")).all()
    g.close()
