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
