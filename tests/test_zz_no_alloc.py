"""fw_reduce_stats_noalloc (the allocation-free variant of fw_reduce_stats,
partials in the library's static device array) against fw_reduce_stats,
through the C ABI: two calls in a row on one GPU, the second on a smaller
domain (so the first call's partials are still in the array), then a domain
taller than the array is refused.  FP64 sums agree to 1e-12 relative (they
are taken in another order), extrema and the wet-cell count exactly."""
import numpy as np
import pytest

from common import load_golden

pytestmark = pytest.mark.gpu


def _check(a, b):
    assert a["wet_cells"] == b["wet_cells"]
    for k in ("max_eta", "min_eta", "max_u", "max_v"):
        assert a[k] == b[k], k
    for k in ("mass_volume", "energy"):
        assert a[k] == pytest.approx(b[k], rel=1e-12, abs=1e-12), k


def test_noalloc_budget_matches_fw_reduce_stats(gpu_available):
    import cases
    import fwcases
    import fwgpu
    import fwsetup
    print("\n[no-alloc budget] 2048 x 2048 beach, then the 64 x 48 beach", flush=True)
    big = fwgpu.Solver(fwsetup.build_state(fwcases.beach_2d(mglob=2048, nglob=2048, dx=2.0), "sp"), "sp")
    big.step(3)
    a, b = big.stats_noalloc(), big.stats()
    _check(a, b)
    assert a["wet_cells"] > 0
    big.close()
    init, _ = load_golden("tiny2d", "dp")
    small = fwgpu.Solver(init, "dp")
    small.step(20)
    a, b = small.stats_noalloc(), small.stats()
    _check(a, b)
    # the same call again gives the same bits
    assert small.stats_noalloc() == a
    small.close()
    st = fwsetup.build_state(cases.ragged2d(), "sp")
    r = fwgpu.Solver(st, "sp")
    r.step(5)
    _check(r.stats_noalloc(), r.stats())
    r.close()


def test_noalloc_refuses_domains_beyond_the_scratch(gpu_available):
    """A domain with more interior rows than the static array holds is
    refused with FW_E_INVALID before any launch."""
    import fwgpu
    init, _ = load_golden("tiny1d", "sp")
    st = dict(init)
    n = 8192 + 1
    st.update(Nglob=n, Nloc=n + 6, Jend=n + 3)
    for k, v in list(st.items()):
        if isinstance(v, np.ndarray) and v.ndim == 2:
            st.pop(k)
    st["PERIODIC"] = 0
    g = fwgpu.Solver(st, "sp", upload=False)
    with pytest.raises(fwgpu.FwError) as e:
        g.stats_noalloc()
    assert e.value.code == fwgpu.FW_E_INVALID
    assert np.isfinite(g.stats()["mass_volume"])
    g.close()
