"""The C restatement (oracle/fw_oracle.c) against the compiled reference.

Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from oracle/_ref) pin whole runs bit for bit; when the compiled reference is
present the restatement is also compared routine by routine on a traced
step, every array, bit for bit."""
import os
import tempfile

import numpy as np
import pytest

from common import STATE_KEYS, bitwise_mismatch, load_golden, oracle_available, ref_available

pytestmark = pytest.mark.skipif(not oracle_available(), reason="oracle not built (oracle/Makefile)")

# every fixture holds steps 1 and 1000 (config0 also its whole 60 s run,
# 1711 steps); the whole-run test checks every step the fixture holds
CASES = ["tiny2d", "tiny1d", "island", "tiny2d_dx15", "config0", "member3", "tiny2d_fsponge",
         "tiny2d_nobrk", "tiny2d_swe", "tiny2d_dy1", "tiny2d_dy15"]


@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("name", CASES)
def test_oracle_whole_run_matches_reference(name, prec):
    import oracle
    init, steps = load_golden(name, prec)
    assert max(steps) >= 1000
    o = oracle.OracleSolver(init)
    done = 0
    for s in sorted(steps):
        rc = o.run(s - done)
        done = s
        assert rc == 0
        ref = steps[s]
        assert o.dt.value == ref["DT"] and o.time.value == ref["TIME"], (s, o.dt.value, ref["DT"])
        bad = {k: bitwise_mismatch(o[k], ref[k]) for k in STATE_KEYS}
        assert all(v == 0 for v in bad.values()), (s, {k: v for k, v in bad.items() if v})


ROUTINES = [("disp", "dispersion"), ("flux", "fluxes"), ("src", "sources"), ("huv", None),
            ("mask", "update_mask"), ("brk", "breaking"), ("exch", "exchange"), ("spg", "sponge")]


@pytest.mark.ref
@pytest.mark.skipif(not ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("prec", ["dp", "sp"])
@pytest.mark.parametrize("name", ["tiny2d", "tiny1d"])
def test_oracle_each_routine_matches_reference(name, prec):
    import cases
    import oracle
    from fwdump import read_dump
    with tempfile.TemporaryDirectory() as td:
        cases.CASES[name]().write(td)
        oracle.run_reference(td, prec, 4, prefix="t", dump_init=False, trace_step=4)
        prev = read_dump(os.path.join(td, "t_trace_0_dt.bin"))
        checked = 0
        for st in (1, 2, 3):
            for tag, meth in ROUTINES:
                f = os.path.join(td, f"t_trace_{st}_{tag}.bin")
                if not os.path.exists(f):
                    continue
                cur = read_dump(f)
                o = oracle.OracleSolver(prev)
                if tag == "huv":
                    o.estimate_huv(st)
                else:
                    getattr(o, meth)()
                for k in oracle.ALL_FIELDS:
                    if isinstance(cur.get(k), np.ndarray) and k not in oracle.AUX:
                        assert bitwise_mismatch(o[k], cur[k]) == 0, (st, tag, k)
                prev = cur
                checked += 1
        assert checked >= 21
