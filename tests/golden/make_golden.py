"""Writes the golden fixtures under tests/golden/ from the compiled reference
(oracle/_ref/fwref_{sp,dp}, built by oracle/Makefile from /root/reference).

For each small workload and precision: the full state after INITIALIZATION
(every array that is not identically zero, plus the scalars) and the key
state arrays after a number of whole time steps, all in full precision.
Every workload reaches step 1000; the reference's 1-D configuration
(config0) also runs to its TOTAL_TIME of 60 s (the loop condition of
src/main.F:284), stored under its step count.
Run from the repo root:  python tests/golden/make_golden.py [case ...]
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cases  # noqa: E402
import oracle  # noqa: E402
from fwdump import read_dump  # noqa: E402

KEY = ["Eta", "U", "V", "Ubar", "Vbar", "HU", "HV", "H", "MASK", "MASK9",
       "AGE_BREAKING", "nu_break", "DepthX", "DepthY", "P", "Q"]
PLAN = {"tiny2d": (["dp", "sp"], [1, 60, 250, 1000]),
        "tiny1d": (["dp", "sp"], [1, 400, 1000]),
        "island": (["dp", "sp"], [1, 120, 1000]),
        "tiny2d_dx15": (["dp", "sp"], [1, 150, 1000]),
        "config0": (["dp", "sp"], [1, 200, 1000, "final"]),
        "member3": (["dp", "sp"], [1, 150, 1000]),
        "tiny2d_fsponge": (["dp", "sp"], [1, 120, 1000]),
        "tiny2d_nobrk": (["dp", "sp"], [1, 120, 1000]),
        "tiny2d_swe": (["dp", "sp"], [1, 120, 1000]),
        "tiny2d_dy1": (["dp", "sp"], [1, 120, 1000]),
        "tiny2d_dy15": (["dp", "sp"], [1, 120, 1000])}
# safety bound on the steps of a run to TOTAL_TIME
MAX_STEPS = 20000


def pack(d, keys=None):
    out = {}
    for k, v in d.items():
        if keys is not None and k not in keys:
            continue
        if isinstance(v, np.ndarray):
            if keys is None and not np.any(v):
                continue
            out[k] = v
        elif isinstance(v, str):
            out[k] = np.array(v)
        else:
            out[k] = np.array(v)
    return out


def main(names=None):
    for name, (precs, steps) in PLAN.items():
        if names and name not in names:
            continue
        for prec in precs:
            with tempfile.TemporaryDirectory() as td:
                c = cases.CASES[name]()
                c.write(td)
                nums = [s for s in steps if s != "final"]
                final = "final" in steps
                out = oracle.run_reference(td, prec, MAX_STEPS if final else max(nums), prefix="g",
                                           dump_init=True, dump_list=nums, until_total=final)
                blob = {}
                for k, v in pack(read_dump(os.path.join(td, "g_init.bin"))).items():
                    blob[f"init/{k}"] = v
                files = [(s, os.path.join(td, f"g_step{s:06d}.bin")) for s in nums]
                if final:  # stored under the step count the run to TOTAL_TIME took
                    nfinal = int([l for l in out.splitlines() if l.startswith("FWREF_STEPS=")][-1].split("=")[1])
                    files.append((nfinal, os.path.join(td, "g_final.bin")))
                for s, f in files:
                    d = read_dump(f)
                    for k, v in pack(d, set(KEY) | {"DT", "TIME"}).items():
                        blob[f"s{s}/{k}"] = v
                path = os.path.join(HERE, f"{name}_{prec}.npz")
                np.savez_compressed(path, **blob)
                print(path, os.path.getsize(path))


if __name__ == "__main__":
    main(sys.argv[1:])
