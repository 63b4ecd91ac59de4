"""The learned coarse-grid correction network: the numpy checker pinned by
brute force and finite differences (CPU), and the HIP kernels against it
(GPU).  The network has no counterpart in the reference; its checker is
tests/cnn_numpy.py.

Tolerances: the device rounds inputs, weights and hidden activations to
bf16 like the checker, but accumulates in fp32 in another order, so an
activation can land on the other side of a bf16 rounding boundary; outputs
agree to 1e-3 relative L2, gradients (carried in bf16 between layers on the
device) to 1e-2 per tensor.
"""
import os
import re

import numpy as np
import pytest

import cnn_numpy as cn
from common import REPO


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def random_params(seed, scale=1.0):
    import fwcnn
    rng = np.random.default_rng(seed)
    p = fwcnn.unflatten(fwcnn.init_params(seed, last_scale=1.0).copy())
    for k in ("b1", "b2", "b3", "b4"):
        p[k][...] = 0.1 * rng.standard_normal(p[k].shape)
    return {k: v * (scale if k.startswith("W") else 1.0) for k, v in p.items()}


def test_parameter_count_matches_header():
    import fwcnn
    src = open(os.path.join(REPO, "include", "fwcnn.h")).read()
    expr = re.search(r"FW_CNN_NPARAMS\s*=\s*([^}]*)\}", src).group(1)
    assert eval(expr) == fwcnn.NPARAMS == 20835


def test_bf16_rounding_is_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5], np.float32)
    exp = np.array([1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -2.5], np.float32)
    np.testing.assert_array_equal(cn.bf16(vals), exp)
    assert cn.bf16(one) == one


def test_numpy_conv_matches_direct_loops():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 5, 7))
    W = rng.standard_normal((9, 3, 4))
    b = rng.standard_normal(4)
    ref = np.zeros((2, 4, 5, 7))
    for bb in range(2):
        for co in range(4):
            for y in range(5):
                for xx in range(7):
                    s = b[co]
                    for t in range(9):
                        dy, dx = t // 3 - 1, t % 3 - 1
                        if 0 <= y + dy < 5 and 0 <= xx + dx < 7:
                            s += sum(x[bb, ci, y + dy, xx + dx] * W[t, ci, co] for ci in range(3))
                    ref[bb, co, y, xx] = s
    np.testing.assert_allclose(cn.conv(x, W, b), ref, rtol=1e-12, atol=1e-12)


def test_numpy_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = {k: np.asarray(v, np.float64) for k, v in random_params(3, 0.7).items()}
    x = rng.standard_normal((2, 5, 6, 5))
    tgt = rng.standard_normal((2, 3, 6, 5))
    _, g = cn.loss_and_grad(p, x, tgt, round_bf16=False)
    h = 1e-6
    for name in ("W1", "b1", "W2", "b3", "W4", "b4"):
        flat = p[name].reshape(-1)
        for k in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            keep = flat[k]
            flat[k] = keep + h
            lp = cn.loss_and_grad(p, x, tgt, round_bf16=False)[0]
            flat[k] = keep - h
            lm = cn.loss_and_grad(p, x, tgt, round_bf16=False)[0]
            flat[k] = keep
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[k]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, k, fd, an)


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_forward_matches_numpy(gpu_available):
    import fwcnn
    rng = np.random.default_rng(4)
    B, H, W = 2, 37, 70  # partial 64-wide tiles and a partial row group
    p = random_params(5)
    x = rng.standard_normal((B, 5, H, W)).astype(np.float32)
    print(f"\n[cnn forward] B={B} H={H} W={W}", flush=True)
    net = fwcnn.Net(H, W, B, fwcnn.flatten(p))
    out = net.forward(x)
    ref, _ = cn.forward(p, x)
    e = rel_l2(out, ref)
    print(f"  rel L2 vs numpy bf16 forward: {e:.2e}")
    assert e < 1e-3
    # batch members are independent
    out1 = net.forward(x[1:2])
    np.testing.assert_array_equal(out1[0], out[1])
    net.close()


@pytest.mark.gpu
def test_gradient_matches_numpy(gpu_available):
    import fwcnn
    rng = np.random.default_rng(6)
    B, H, W = 2, 20, 66
    p = random_params(7)
    x = rng.standard_normal((B, 5, H, W)).astype(np.float32)
    tgt = rng.standard_normal((B, 3, H, W)).astype(np.float32)
    print(f"\n[cnn gradient] B={B} H={H} W={W}", flush=True)
    net = fwcnn.Net(H, W, B, fwcnn.flatten(p))
    loss, g = net.train_step(x, tgt, lr=0.0, want_grad=True)
    ref_loss, ref_g = cn.loss_and_grad(p, x, tgt)
    assert loss == pytest.approx(ref_loss, rel=1e-2)
    gd = fwcnn.unflatten(g)
    for name, _ in fwcnn.SHAPES:
        e = rel_l2(gd[name], ref_g[name])
        print(f"  {name}: rel L2 {e:.2e}")
        assert e < 1e-2, name
    # lr = 0 leaves the parameters unchanged
    np.testing.assert_array_equal(net.params(), fwcnn.flatten(p))
    net.close()


@pytest.mark.gpu
def test_training_reduces_loss(gpu_available):
    """Fit a smooth function of the inputs: the loss falls tenfold."""
    import fwcnn
    rng = np.random.default_rng(8)
    B, H, W = 4, 32, 64
    net = fwcnn.Net(H, W, B, fwcnn.init_params(9))
    x = rng.standard_normal((B, 5, H, W)).astype(np.float32)
    tgt = np.stack([0.5 * x[:, 0], 0.3 * x[:, 1] - 0.2 * x[:, 2], 0.1 * x[:, 3]], axis=1).astype(np.float32)
    losses = [net.train_step(x, tgt, lr=3e-3)[0] for _ in range(300)]
    print(f"\n[cnn training] loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.1 * losses[0]
    net.close()


@pytest.mark.gpu
def test_correction_updates_wet_cells(gpu_available):
    """With all weights zero the network outputs its last bias: the
    correction adds bias * output scale on wet interior cells only and keeps
    H, HU, HV, Ubar, Vbar consistent with the new eta, u, v."""
    import fwcnn
    import fwgpu
    from common import load_golden
    init, _ = load_golden("tiny2d", "sp")
    g = fwgpu.Solver(init, "sp")
    g.step(2)
    before = {k: g.download(k).astype(np.float64) for k in ("Eta", "U", "V", "H", "HU", "HV", "Ubar", "Vbar",
                                                            "Depth", "MASK")}
    H, W = init["Nloc"] - 6, init["Mloc"] - 6
    p = np.zeros(fwcnn.NPARAMS, np.float32)
    d = fwcnn.unflatten(p)
    d["b4"][...] = [2.0, -1.0, 0.5]
    net = fwcnn.Net(H, W, 1, fwcnn.flatten(d))
    os_ = np.array([1e-3, 2e-3, 4e-3], np.float32)
    net.correct(g, os_=os_)
    g.synchronize()
    after = {k: g.download(k).astype(np.float64) for k in before}
    inner = (slice(3, -3), slice(3, -3))
    wet = before["MASK"][inner] == 1
    de = np.float32(np.float32(2.0) * os_[0])
    np.testing.assert_array_equal(after["Eta"][inner][wet], (before["Eta"][inner][wet].astype(np.float32) + de))
    np.testing.assert_array_equal(after["Eta"][inner][~wet], before["Eta"][inner][~wet])
    hm = np.maximum(after["H"][inner][wet], init["MinDepthFrc"])
    np.testing.assert_allclose(after["HU"][inner][wet], hm * after["U"][inner][wet], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(after["Ubar"][inner][wet] - before["Ubar"][inner][wet],
                               hm * (after["U"][inner][wet] - before["U"][inner][wet]), rtol=1e-4, atol=1e-6)
    # ghost rows untouched
    np.testing.assert_array_equal(after["Eta"][:3], before["Eta"][:3])
    net.close()
    g.close()


@pytest.mark.gpu
def test_conserving_correction_moves_no_water(gpu_available):
    """With set_conserve the wet-cell mean of the eta correction is taken
    out: a constant network output leaves eta unchanged (u, v still move),
    and a random network changes the wet-cell sum of eta only by rounding."""
    import fwcnn
    import fwgpu
    from common import load_golden
    print("\n[conserving correction] tiny2d sp", flush=True)
    init, _ = load_golden("tiny2d", "sp")
    g = fwgpu.Solver(init, "sp")
    g.step(2)
    H, W = init["Nloc"] - 6, init["Mloc"] - 6
    inner = (slice(3, -3), slice(3, -3))
    wet = g.download("MASK")[inner] == 1
    os_ = np.array([1e-3, 2e-3, 4e-3], np.float32)
    d = fwcnn.unflatten(np.zeros(fwcnn.NPARAMS, np.float32))
    d["b4"][...] = [2.0, -1.0, 0.5]
    net = fwcnn.Net(H, W, 1, fwcnn.flatten(d))
    net.set_conserve(True)
    e0, u0 = g.download("Eta")[inner], g.download("U")[inner]
    net.correct(g, os_=os_)
    g.synchronize()
    e1, u1 = g.download("Eta")[inner], g.download("U")[inner]
    np.testing.assert_array_equal(e1, e0)
    assert np.all(u1[wet] != u0[wet])
    net.close()
    net = fwcnn.Net(H, W, 1, fwcnn.init_params(seed=3, last_scale=1.0))
    net.set_conserve(True)
    e0 = g.download("Eta")[inner].astype(np.float64)
    net.correct(g, os_=os_)
    g.synchronize()
    e1 = g.download("Eta")[inner].astype(np.float64)
    moved = np.sum(np.abs(e1 - e0)[wet])
    assert moved > 0
    assert abs(np.sum((e1 - e0)[wet])) <= 1e-3 * moved
    net.close()
    g.close()
