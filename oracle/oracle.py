"""ctypes front end of the C restatement (fw_oracle.c) and of the compiled
reference (oracle/_ref/fwref_{sp,dp}).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's cpu_baseline
leg, as the checker.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_HDR = open(os.path.join(HERE, "fw_oracle.h")).read()


def _macro_names(macro: str) -> list[str]:
    body = _HDR.split(f"#define {macro}(X)")[1].split("\n#")[0]
    import re
    return re.findall(r"X\((\w+)\)", body)


CELL = _macro_names("FWO_CELL_FIELDS")
XFACE = _macro_names("FWO_XFACE_FIELDS")
YFACE = _macro_names("FWO_YFACE_FIELDS")
INTF = _macro_names("FWO_INT_FIELDS")
AUX = _macro_names("FWO_AUX_FIELDS")
ALL_FIELDS = CELL + XFACE + YFACE + INTF + AUX

INT_PARAMS = ["Mloc", "Nloc", "Mglob", "Nglob", "Ibeg", "Iend", "Jbeg", "Jend",
              "DISPERSION", "PERIODIC", "VISCOSITY_BREAKING", "SHOW_BREAKING",
              "DIFFUSION_SPONGE", "FRICTION_SPONGE", "DIRECT_SPONGE",
              "ETA_LIMITER", "FIXED_DT", "wavemaker", "NumFreq", "ilo", "ihi",
              "jlo", "jhi"]
DBL_PARAMS = ["DX", "DY", "DXg", "DYg", "CFL", "MinDepth", "MinDepthFrc",
              "Gamma1", "Gamma2", "Gamma3", "Beta_1", "Beta_2", "b1", "b2",
              "SWE_ETA_DEP", "FroudeCap", "Cbrk1", "Cbrk2", "T_brk",
              "WAVEMAKER_Cbrk", "nu_bkg", "Xc_WK", "Yc_WK", "Width_WK",
              "Ywidth_WK", "EtaBlowVal", "DT_fixed", "CrestLimit",
              "TroughLimit", "D_gen", "Beta_gen", "rlamda", "Tperiod",
              "Time_ramp", "FreqPeak"]


class Params(ct.Structure):
    _fields_ = [(n, ct.c_int) for n in INT_PARAMS] + [(n, ct.c_double) for n in DBL_PARAMS]


class Fields(ct.Structure):
    _fields_ = [(n, ct.c_void_p) for n in ALL_FIELDS]


WAVEMAKERS = {"WK_REG": 1, "WK_DATA2D": 2}
# wavemakers whose per-step path needs no source evaluation
NO_SOURCE_WAVEMAKERS = ("INI_SOL", "INI_REC", "INI_GAU", "INI_OTH", "N_WAVE", "NONE")


def wavemaker_code(name: str) -> int:
    name = name.strip()
    for k, v in WAVEMAKERS.items():
        if name.startswith(k):
            return v
    if any(name.startswith(k) for k in NO_SOURCE_WAVEMAKERS) or name == "":
        return 0
    raise ValueError(f"wavemaker {name!r} is not on the supported per-step path")


def params_from_dump(d: dict) -> Params:
    p = Params()
    for n in INT_PARAMS:
        if n == "wavemaker":
            p.wavemaker = wavemaker_code(d["WaveMaker"])
        elif n in d:
            setattr(p, n, int(d[n]))
    for n in DBL_PARAMS:
        if n in d:
            setattr(p, n, float(d[n]))
    return p


def load_lib(precision: str):
    lib = ct.CDLL(os.path.join(REF_DIR, f"libfworacle_{precision}.so"))
    P, F = ct.POINTER(Params), ct.POINTER(Fields)
    dp = ct.POINTER(ct.c_double)
    lib.fwo_exchange.argtypes = [P, F]
    lib.fwo_estimate_dt.argtypes = [P, F, dp, dp]
    lib.fwo_dispersion.argtypes = [P, F, ct.c_double]
    lib.fwo_fluxes.argtypes = [P, F]
    lib.fwo_sources.argtypes = [P, F, ct.c_double]
    lib.fwo_estimate_huv.argtypes = [P, F, ct.c_int, ct.c_double]
    lib.fwo_update_mask.argtypes = [P, F]
    lib.fwo_breaking.argtypes = [P, F, ct.c_double]
    lib.fwo_sponge.argtypes = [P, F]
    lib.fwo_max_abs_eta.argtypes = [P, F]
    lib.fwo_max_abs_eta.restype = ct.c_double
    lib.fwo_run.argtypes = [P, F, ct.c_int, dp, dp]
    lib.fwo_run.restype = ct.c_int
    return lib


class OracleSolver:
    """Holds a full copy of the state (numpy, [j, i] = Fortran (I, J)) and
    advances it with the C restatement."""

    def __init__(self, dump: dict):
        self.precision = "dp" if dump["SP_BYTES"] == 8 else "sp"
        self.rdt = np.float64 if self.precision == "dp" else np.float32
        self.lib = load_lib(self.precision)
        self.p = params_from_dump(dump)
        M, N = self.p.Mloc, self.p.Nloc
        self.arrays = {}
        shapes = {}
        for n in CELL:
            shapes[n] = (N, M)
        for n in XFACE:
            shapes[n] = (N, M + 1)
        for n in YFACE:
            shapes[n] = (N + 1, M)
        for n in INTF:
            shapes[n] = (N, M)
        for n in ALL_FIELDS:
            if n in dump and isinstance(dump[n], np.ndarray):
                a = np.ascontiguousarray(dump[n])
                if n in INTF:
                    a = a.astype(np.int32)
                else:
                    a = a.astype(self.rdt)
                self.arrays[n] = a.copy()
            elif n in shapes:
                dt = np.int32 if n in INTF else self.rdt
                self.arrays[n] = np.zeros(shapes[n], dt)
            else:
                self.arrays[n] = np.zeros(1, self.rdt)
        self.f = Fields()
        for n in ALL_FIELDS:
            setattr(self.f, n, self.arrays[n].ctypes.data)
        self.dt = ct.c_double(float(dump.get("DT", 0.0)))
        self.time = ct.c_double(float(dump.get("TIME", 0.0)))

    def __getitem__(self, k):
        return self.arrays[k]

    def run(self, nsteps: int) -> int:
        return self.lib.fwo_run(ct.byref(self.p), ct.byref(self.f), nsteps,
                                ct.byref(self.dt), ct.byref(self.time))

    # single routines, same names as the reference
    def exchange(self):
        self.lib.fwo_exchange(ct.byref(self.p), ct.byref(self.f))

    def estimate_dt(self):
        self.lib.fwo_estimate_dt(ct.byref(self.p), ct.byref(self.f),
                                 ct.byref(self.dt), ct.byref(self.time))

    def dispersion(self):
        self.lib.fwo_dispersion(ct.byref(self.p), ct.byref(self.f), self.dt.value)

    def fluxes(self):
        self.lib.fwo_fluxes(ct.byref(self.p), ct.byref(self.f))

    def sources(self):
        self.lib.fwo_sources(ct.byref(self.p), ct.byref(self.f), self.time.value)

    def estimate_huv(self, istage):
        self.lib.fwo_estimate_huv(ct.byref(self.p), ct.byref(self.f), istage, self.dt.value)

    def update_mask(self):
        self.lib.fwo_update_mask(ct.byref(self.p), ct.byref(self.f))

    def breaking(self):
        self.lib.fwo_breaking(ct.byref(self.p), ct.byref(self.f), self.dt.value)

    def sponge(self):
        self.lib.fwo_sponge(ct.byref(self.p), ct.byref(self.f))


def ref_binary(precision: str) -> str:
    return os.path.join(REF_DIR, f"fwref_{precision}")


def run_reference(case_dir: str, precision: str, nsteps: int, prefix="d",
                  dump_init=True, dump_every=0, trace_step=-1, warmup=-1,
                  output=False, timeout=600, dump_list=(), until_total=False):
    """Run the compiled reference in ``case_dir`` (containing input.txt).
    ``dump_list``: steps after which the state is dumped; ``until_total``:
    also stop once TIME >= TOTAL_TIME (src/main.F:284) and dump the final
    state as <prefix>_final.bin."""
    env = dict(os.environ)
    env.update({"FWREF_PREFIX": prefix, "FWREF_NSTEPS": str(nsteps),
                "FWREF_DUMP_INIT": "1" if dump_init else "0",
                "FWREF_DUMP_EVERY": str(dump_every),
                "FWREF_DUMP_LIST": ",".join(str(int(s)) for s in dump_list),
                "FWREF_UNTIL_TOTAL": "1" if until_total else "0",
                "FWREF_TRACE_STEP": str(trace_step),
                "FWREF_WARMUP": str(warmup),
                "FWREF_OUTPUT": "1" if output else "0",
                # fresh heap allocations read as zero: the reference reads an
                # uninitialised MASK9tmp ring (masks.F:182)
                "MALLOC_PERTURB_": "255"})
    cmd = f"ulimit -s unlimited; exec {ref_binary(precision)} input.txt"
    r = subprocess.run(["bash", "-c", cmd], cwd=case_dir, env=env,
                       capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"reference run failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r.stdout
