"""WAVE_BREAKING (breaker.F:60-353) with the roller flux and undertow of
continuing breaking (:326-337), pinned by the compiled reference's own
arrays at step 250, stage 3 of tiny2d (tests/golden/make_breaking_golden.py):
the oracle bit for bit on CPU, the GPU against the same fixture."""
import numpy as np
import pytest

from common import bitwise_mismatch, load_breaking_golden, oracle_available, ulp_distance

POST = ["AGE_BREAKING", "nu_break", "ROLLER_FLUX", "UNDERTOW_U", "UNDERTOW_V"]


@pytest.mark.skipif(not oracle_available(), reason="oracle not built")
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_oracle_breaking_matches_reference(prec):
    import oracle
    pre, post = load_breaking_golden(prec)
    assert np.count_nonzero(post["ROLLER_FLUX"]) > 100  # continuing breaking is exercised
    assert np.any(post["ROLLER_FLUX"] != pre["ROLLER_FLUX"])
    o = oracle.OracleSolver(pre)
    o.breaking()
    for k in POST:
        assert bitwise_mismatch(o[k], post[k]) == 0, k


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_gpu_breaking_matches_reference(prec, gpu_available):
    """AGE_BREAKING, nu_break and ROLLER_FLUX bit for bit; the undertow
    multiplies the roller flux by COS/SIN of the breaking angle, which the
    device math library evaluates (correctly rounded in double, then
    rounded once in FP32) where the reference calls the host's: within
    2 ulp.  The breaking angle itself comes from ATAN2 the same way."""
    import fwgpu
    print(f"\n[breaking golden] prec={prec}", flush=True)
    pre, post = load_breaking_golden(prec)
    g = fwgpu.Solver(pre, prec, full_state=True)
    g.breaking()
    for k in ("AGE_BREAKING", "nu_break", "ROLLER_FLUX"):
        assert bitwise_mismatch(g.download(k), post[k]) == 0, k
    for k in ("UNDERTOW_U", "UNDERTOW_V"):
        d = ulp_distance(g.download(k), post[k])
        print(f"{k}: max {d} ulp", flush=True)
        assert d <= 2, (k, d)
    g.close()


@pytest.mark.gpu
@pytest.mark.skipif(not oracle_available(), reason="oracle not built")
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_gpu_roller_uses_mean_surface(prec, gpu_available):
    """The roller flux scales with eta - ETAmean (ETAmean is zero in the
    reference run unless mean statistics are kept): a random ETAmean,
    GPU against the oracle."""
    import fwgpu
    import oracle
    print(f"\n[roller ETAmean] prec={prec}", flush=True)
    pre, _ = load_breaking_golden(prec)
    rng = np.random.default_rng(3)
    pre = dict(pre)
    pre["ETAmean"] = (rng.normal(size=pre["Eta"].shape) * 0.1).astype(pre["Eta"].dtype)
    g, o = fwgpu.Solver(pre, prec, full_state=True), oracle.OracleSolver(pre)
    g.breaking(); o.breaking()
    assert bitwise_mismatch(g.download("ROLLER_FLUX"), o["ROLLER_FLUX"]) == 0
    assert np.count_nonzero(o["ROLLER_FLUX"]) > 100
    for k in ("UNDERTOW_U", "UNDERTOW_V"):
        assert ulp_distance(g.download(k), o[k]) <= 2, k
    g.close()


@pytest.mark.gpu
@pytest.mark.skipif(not oracle_available(), reason="oracle not built")
@pytest.mark.parametrize("prec", ["sp", "dp"])
def test_fused_step_roller_matches_oracle(prec, gpu_available):
    """fw_step with every intermediate kept, from tiny2d's step-250 state
    (breaking under way), 5 steps: the roller flux and undertow against the
    oracle's whole steps.  In fw_step the breaking kernels run beside the
    line solves; the roller must read U and V after them.  Same cells, and
    within the whole-run tolerance (1e-5 / 1e-10 relative L2: the steps
    evaluate ATAN2/COS/SIN per cell)."""
    import fwgpu
    import oracle
    from common import load_golden, rel_l2
    print(f"\n[fused roller] prec={prec}", flush=True)
    init, steps = load_golden("tiny2d", prec)
    st = dict(init)
    st.update({k: v for k, v in steps[250].items() if not np.isscalar(v)})
    st["DT"], st["TIME"] = steps[250]["DT"], steps[250]["TIME"]
    g, o = fwgpu.Solver(st, prec, full_state=True), oracle.OracleSolver(st)
    g.step(5)
    o.run(5)
    tol = 1e-5 if prec == "sp" else 1e-10
    assert np.count_nonzero(o["ROLLER_FLUX"]) > 0
    for k in ("ROLLER_FLUX", "UNDERTOW_U", "UNDERTOW_V"):
        a, b = g.download(k), o[k]
        assert np.array_equal(a != 0, b != 0), k
        assert rel_l2(a, b) <= tol, (k, rel_l2(a, b))
    g.close()
