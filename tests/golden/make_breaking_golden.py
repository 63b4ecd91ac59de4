"""Writes tests/golden/breaking_tiny2d_{sp,dp}.npz from the compiled
reference (oracle/_ref/fwref_{sp,dp}, built by oracle/Makefile from
/root/reference): the state WAVE_BREAKING reads at step 250, stage 3 of the
tiny2d workload (after UPDATE_MASK; continuing breaking is under way there)
and the arrays it writes (AGE_BREAKING, nu_break and the roller flux and
undertow of breaker.F:326-337).  Run from the repo root:
python tests/golden/make_breaking_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import cases  # noqa: E402
import oracle  # noqa: E402
from fwdump import read_dump  # noqa: E402

STEP = 250
PRE = ["MASK", "MASK9", "Eta", "U", "V", "H", "Depth", "AGE_BREAKING", "nu_break", "ETAT", "ETAx", "ETAy",
       "ROLLER_FLUX", "UNDERTOW_U", "UNDERTOW_V", "ETAmean", "xmk_wk", "ymk_wk"]
POST = ["AGE_BREAKING", "nu_break", "ROLLER_FLUX", "UNDERTOW_U", "UNDERTOW_V"]


def main():
    for prec in ("sp", "dp"):
        with tempfile.TemporaryDirectory() as td:
            cases.CASES["tiny2d"]().write(td)
            oracle.run_reference(td, prec, STEP, prefix="t", dump_init=False, trace_step=STEP)
            pre = read_dump(os.path.join(td, "t_trace_3_mask.bin"))
            post = read_dump(os.path.join(td, "t_trace_3_brk.bin"))
        out = {}
        for k, v in pre.items():
            if isinstance(v, np.ndarray):
                if k in PRE and np.any(v):
                    out["pre/" + k] = v
            else:
                out["pre/" + k] = np.array(v)
        for k in POST:
            out["post/" + k] = post[k]
        path = os.path.join(HERE, f"breaking_tiny2d_{prec}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path), "bytes;", int(np.count_nonzero(post["ROLLER_FLUX"])), "roller cells")


if __name__ == "__main__":
    main()
