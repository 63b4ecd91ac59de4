"""Learned coarse-grid ensembles (funwave-tvd_amd/fwlearn.py): grid
coarsening on the host (CPU) and the whole train-and-compare workflow on a
small member (GPU)."""
import numpy as np
import pytest


def test_coarse_member_is_the_block_mean():
    import fwlearn
    c = fwlearn.member(3, mglob=40, nglob=24, dt=0.1, r_base=20.0, h0=2.0)
    cc = fwlearn.coarse_case(c)
    assert (cc.mglob, cc.nglob) == (20, 12)
    assert cc.params["DX"] == 2 * c.params["DX"] and cc.params["DT_fixed"] == pytest.approx(0.2)
    d = np.asarray(c.depth)
    np.testing.assert_allclose(cc.depth[1, 2], d[2:4, 4:6].mean())
    np.testing.assert_allclose(fwlearn.restrict(d).sum() * 4, d.sum())
    assert c.params["WAVEMAKER"] == "INI_SOL"


def test_fixed_step_reaches_the_host_state():
    import fwlearn
    import fwsetup
    c = fwlearn.member(4, mglob=40, nglob=24, r_base=20.0, h0=2.0)
    dt = fwlearn.stable_dt(c)
    st = fwsetup.build_state(fwlearn.with_fixed_dt(c, dt), "sp")
    assert st["FIXED_DT"] == 1 and st["DT_fixed"] == dt
    assert 0.0 < dt < 0.5 * c.params["DX"] / np.sqrt(9.81 * np.max(c.depth))


@pytest.mark.gpu
def test_learned_workflow_end_to_end(gpu_available):
    """Training data from sixteen members (192 x 128 fine cells), a network
    trained on them with the project's own kernels, and four members of
    another size (96 x 64) that the network has not seen: on their one-step
    samples the corrected coarse step must leave less error than the plain
    coarse step, and in the batched comparison over 40 fine steps the
    corrected coarse runs must end closer to the restricted fine solution
    than the plain coarse runs, on average over the four (what the learned
    part gains over the method without it), with finite, positive
    runs/hour."""
    import fwcnn
    import fwlearn
    print("\n[learned workflow] members 20..35 (192 x 128) train; members 3, 6, 13, 14 (96 x 64) held out",
          flush=True)

    def case(seed, m, n, r):
        base = fwlearn.member(seed, mglob=m, nglob=n, r_base=r)
        return fwlearn.with_fixed_dt(base, fwlearn.stable_dt(base))

    data = [fwlearn.make_samples(case(s, 192, 128, 80.0), n_samples=16, every=2) for s in range(20, 36)]
    x, y = data[0]
    assert x.shape == (16, 5, 64, 96) and y.shape == (16, 3, 64, 96)
    assert np.isfinite(x).all() and np.isfinite(y).all()
    params, losses = fwlearn.train(data, steps=6000, batch=16, size=32, lr=2e-3)
    print(f"  loss {np.mean(losses[:20]):.3e} -> {np.mean(losses[-20:]):.3e}")
    assert np.mean(losses[-20:]) < np.mean(losses[:20])
    held = [case(s, 96, 64, 40.0) for s in (3, 6, 13, 14)]
    # one-step: the held-out members' samples, corrected vs uncorrected
    net = fwcnn.Net(32, 48, 12, params)
    e_net, e_zero = 0.0, 0.0
    for c in held:
        hx, hy = fwlearn.make_samples(c, n_samples=12, every=2)
        e_net += float(np.sum((net.forward(hx) - hy) ** 2))
        e_zero += float(np.sum(hy ** 2))
    net.close()
    print(f"  held-out one-step squared error: corrected {e_net:.4e}, uncorrected {e_zero:.4e}")
    assert e_net < e_zero
    r = fwlearn.compare_batched(held, fine_steps=40, params=params)
    for row in r["members"]:
        print(f"  held-out {row['member']}: coarse error {row['err_coarse']:.3e}, "
              f"learned {row['err_learned']:.3e}")
    print(f"  runs/h {r['runs_per_hour']}")
    errs_l = [row["err_learned"] for row in r["members"]]
    errs_c = [row["err_coarse"] for row in r["members"]]
    assert np.isfinite(errs_l).all()
    assert np.mean(errs_l) < np.mean(errs_c)
    assert all(v > 0 for v in r["runs_per_hour"].values())
