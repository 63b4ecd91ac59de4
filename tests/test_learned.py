"""Learned coarse-grid ensembles (funwave-tvd_amd/fwlearn.py): grid
coarsening on the host (CPU) and the whole train-and-compare workflow on a
small member (GPU)."""
import numpy as np
import pytest


def test_coarse_member_is_the_block_mean():
    import fwlearn
    c = fwlearn.member(3, mglob=40, nglob=24, dt=0.1)
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
    c = fwlearn.member(4, mglob=40, nglob=24)
    dt = fwlearn.stable_dt(c)
    st = fwsetup.build_state(fwlearn.with_fixed_dt(c, dt), "sp")
    assert st["FIXED_DT"] == 1 and st["DT_fixed"] == dt
    assert 0.0 < dt < 0.5 * c.params["DX"] / np.sqrt(9.81 * np.max(c.depth))


@pytest.mark.gpu
def test_learned_workflow_end_to_end(gpu_available):
    import fwlearn
    print("\n[learned workflow] member 1, 96 x 64 fine cells", flush=True)
    base = fwlearn.member(1, mglob=96, nglob=64)
    case = fwlearn.with_fixed_dt(base, fwlearn.stable_dt(base))
    x, y = fwlearn.make_samples(case, n_samples=6, every=2)
    assert x.shape == (6, 5, 32, 48) and y.shape == (6, 3, 32, 48)
    assert np.isfinite(x).all() and np.isfinite(y).all()
    params, losses = fwlearn.train([(x, y)], steps=150, batch=8, size=32, lr=2e-3)
    print(f"  loss {np.mean(losses[:10]):.3e} -> {np.mean(losses[-10:]):.3e}")
    assert np.mean(losses[-10:]) < np.mean(losses[:10])
    r = fwlearn.compare([case], fine_steps=40, params=params)
    row = r["members"][0]
    print(f"  errors vs restricted fine: coarse {row['err_coarse']:.3e}, learned {row['err_learned']:.3e}; "
          f"runs/h {r['runs_per_hour']}")
    assert np.isfinite(row["err_learned"]) and row["err_learned"] < 1.0
    assert all(v > 0 for v in r["runs_per_hour"].values())
