"""Shared helpers of the parity tests."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
REPO = os.path.dirname(HERE)

# arrays compared after whole steps
STATE_KEYS = ["Eta", "U", "V", "Ubar", "Vbar", "HU", "HV", "H", "MASK", "MASK9",
              "AGE_BREAKING", "nu_break", "DepthX", "DepthY", "P", "Q"]


def load_golden(name, prec):
    """(init_state, {step: arrays}) from tests/golden/<name>_<prec>.npz."""
    z = np.load(os.path.join(GOLDEN, f"{name}_{prec}.npz"), allow_pickle=False)
    init, steps = {}, {}
    for k in z.files:
        grp, key = k.split("/", 1)
        v = z[k]
        if v.ndim == 0:
            v = v.item()
        if grp == "init":
            init[key] = v
        else:
            steps.setdefault(int(grp[1:]), {})[key] = v
    return init, steps


def ref_available():
    return all(os.path.exists(os.path.join(REPO, "oracle", "_ref", f)) for f in ("fwref_sp", "fwref_dp"))


def oracle_available():
    return all(os.path.exists(os.path.join(REPO, "oracle", "_ref", f))
               for f in ("libfworacle_sp.so", "libfworacle_dp.so"))


def bitwise_mismatch(a, b):
    """number of elements whose bits differ (NaN == NaN)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return -1
    if a.dtype.kind == "f":
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        # +0 and -0 compare equal numerically; require the same sign too
        same &= np.signbit(a) == np.signbit(b)
    else:
        same = a == b
    return int((~same).sum())


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return float(d / n) if n > 0 else float(d)


def load_breaking_golden(prec):
    """(state WAVE_BREAKING reads, arrays it writes) at step 250, stage 3 of
    tiny2d, from tests/golden/breaking_tiny2d_<prec>.npz."""
    z = np.load(os.path.join(GOLDEN, f"breaking_tiny2d_{prec}.npz"), allow_pickle=False)
    pre, post = {}, {}
    for k in z.files:
        grp, key = k.split("/", 1)
        v = z[k]
        if v.ndim == 0:
            v = v.item()
        (pre if grp == "pre" else post)[key] = v
    return pre, post


def ulp_distance(a, b):
    """Largest distance in units in the last place between two arrays of
    the same float dtype (0 for identical bits)."""
    a, b = np.asarray(a), np.asarray(b)
    it = np.int64 if a.dtype == np.float64 else np.int32
    ia, ib = a.view(it).astype(np.int64), b.view(it).astype(np.int64)
    # map the sign-magnitude bit patterns onto a monotone integer line
    top = np.int64(1) << (63 if it is np.int64 else 31)
    ia = np.where(ia < 0, -(ia + top) if it is np.int64 else -(ia & 0x7fffffff), ia)
    ib = np.where(ib < 0, -(ib + top) if it is np.int64 else -(ib & 0x7fffffff), ib)
    return int(np.max(np.abs(ia - ib))) if a.size else 0
