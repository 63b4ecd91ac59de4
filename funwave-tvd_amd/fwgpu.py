"""Python host binding of libfwgpu.so (the C ABI in include/fwgpu.h).

Mirrors the reference's per-step interface: one method per routine the
reference main loop calls (src/main.F:295-517), same names and order, plus
whole-step driving, state transfer in the reference array layout and
measurement hooks.  The product path is the HIP library only: if it cannot
be loaded this module raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfwgpu.so")

INT_CFG = ["real_bytes", "Mloc", "Nloc", "Mglob", "Nglob", "Ibeg", "Iend", "Jbeg",
           "Jend", "DISPERSION", "PERIODIC", "VISCOSITY_BREAKING", "SHOW_BREAKING",
           "DIFFUSION_SPONGE", "FRICTION_SPONGE", "DIRECT_SPONGE", "ETA_LIMITER",
           "FIXED_DT", "wavemaker", "NumFreq", "ilo", "ihi", "jlo", "jhi"]
DBL_CFG = ["DX", "DY", "DXg", "DYg", "CFL", "MinDepth", "MinDepthFrc", "Gamma1",
           "Gamma2", "Gamma3", "Beta_1", "Beta_2", "b1", "b2", "SWE_ETA_DEP",
           "FroudeCap", "Cbrk1", "Cbrk2", "T_brk", "WAVEMAKER_Cbrk", "nu_bkg",
           "Xc_WK", "Yc_WK", "Width_WK", "Ywidth_WK", "EtaBlowVal", "DT_fixed",
           "CrestLimit", "TroughLimit", "D_gen", "Beta_gen", "rlamda", "Tperiod",
           "Time_ramp", "FreqPeak"]

FW_OPT_FULL_STATE = 1
FW_OPT_HOST_MATH = 2
FW_OK, FW_E_INVALID, FW_E_HIP, FW_E_NOMEM, FW_E_BLOWUP, FW_E_NOTFOUND, FW_E_BREAK_ITER = 0, -1, -2, -3, -4, -5, -6
WK = {"WK_REG": 1, "WK_DATA2D": 2}
NO_SOURCE_WAVEMAKERS = ("INI_SOL", "INI_REC", "INI_GAU", "INI_OTH", "N_WAVE", "NONE")

# exported symbols (every one is declared in include/fwgpu.h)
EXPORTS = ["fw_create", "fw_destroy", "fw_strerror", "fw_synchronize",
           "fw_device_bytes", "fw_field_id", "fw_upload", "fw_download",
           "fw_set_clock", "fw_get_clock", "fw_exchange", "fw_estimate_dt",
           "fw_save_stage0", "fw_cal_dispersion", "fw_fluxes", "fw_source_terms",
           "fw_estimate_huv", "fw_update_mask", "fw_wave_breaking",
           "fw_sponge_damping", "fw_check_blowup", "fw_step", "fw_reduce_stats",
           "fw_set_option", "fw_profile_enable", "fw_profile_reset", "fw_profile_get",
           "fw_profile_names", "fw_time_steps", "fw_field_device", "fw_get_stream",
           "fw_get_config", "fw_batch_create", "fw_batch_destroy", "fw_batch_strerror",
           "fw_batch_size", "fw_batch_step", "fw_batch_status", "fw_batch_time_steps",
           "fw_batch_set_graph", "fw_batch_autotune", "fw_batch_get_shapes", "fw_reduce_stats_noalloc"]
FW_BATCH_RUNNING, FW_BATCH_BLOWUP, FW_BATCH_DT_CHANGED, FW_BATCH_BREAK_ITER = 0, 1, 2, 3


class Config(ct.Structure):
    _fields_ = [(n, ct.c_int) for n in INT_CFG] + [(n, ct.c_double) for n in DBL_CFG]


class Stats(ct.Structure):
    _fields_ = [("mass_volume", ct.c_double), ("energy", ct.c_double),
                ("max_eta", ct.c_double), ("min_eta", ct.c_double),
                ("max_u", ct.c_double), ("max_v", ct.c_double),
                ("wet_cells", ct.c_int64)]


class FwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fwgpu error {code}: {msg}")
        self.code = code


_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libfwgpu.so (raises if absent: there is no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built; run __graft_entry__.build()")
    lib = ct.CDLL(path)
    vp, ip, dp = ct.c_void_p, ct.c_int, ct.POINTER(ct.c_double)
    lib.fw_create.argtypes = [ct.POINTER(Config), ct.POINTER(vp)]
    lib.fw_destroy.argtypes = [vp]
    lib.fw_strerror.argtypes = [vp]
    lib.fw_strerror.restype = ct.c_char_p
    lib.fw_synchronize.argtypes = [vp]
    lib.fw_device_bytes.argtypes = [vp, ct.POINTER(ct.c_int64)]
    lib.fw_field_id.argtypes = [ct.c_char_p]
    lib.fw_upload.argtypes = [vp, ip, vp]
    lib.fw_download.argtypes = [vp, ip, vp]
    lib.fw_set_clock.argtypes = [vp, ct.c_double, ct.c_double]
    lib.fw_get_clock.argtypes = [vp, dp, dp]
    for n in ["fw_exchange", "fw_cal_dispersion", "fw_fluxes", "fw_source_terms",
              "fw_update_mask", "fw_wave_breaking", "fw_sponge_damping",
              "fw_profile_reset"]:
        getattr(lib, n).argtypes = [vp]
    lib.fw_estimate_dt.argtypes = [vp, dp, dp]
    lib.fw_save_stage0.argtypes = [vp, ip]
    lib.fw_estimate_huv.argtypes = [vp, ip]
    lib.fw_check_blowup.argtypes = [vp, dp]
    lib.fw_step.argtypes = [vp, ip, dp, dp]
    lib.fw_reduce_stats.argtypes = [vp, ct.POINTER(Stats)]
    lib.fw_reduce_stats_noalloc.argtypes = [vp, ct.POINTER(Stats)]
    lib.fw_profile_enable.argtypes = [vp, ip]
    lib.fw_set_option.argtypes = [vp, ip, ip]
    lib.fw_profile_get.argtypes = [vp, ct.c_char_p, ct.POINTER(ct.c_int64), dp]
    lib.fw_profile_names.argtypes = [vp, ct.c_char_p, ip]
    lib.fw_time_steps.argtypes = [vp, ip, dp]
    lib.fw_batch_create.argtypes = [ct.POINTER(vp), ip, ct.POINTER(vp)]
    lib.fw_batch_destroy.argtypes = [vp]
    lib.fw_batch_strerror.argtypes = [vp]
    lib.fw_batch_strerror.restype = ct.c_char_p
    lib.fw_batch_size.argtypes = [vp, ct.POINTER(ip)]
    lib.fw_batch_step.argtypes = [vp, ip]
    lib.fw_batch_status.argtypes = [vp, ct.POINTER(ip)]
    lib.fw_batch_time_steps.argtypes = [vp, ip, dp]
    lib.fw_batch_set_graph.argtypes = [vp, ip]
    lib.fw_batch_autotune.argtypes = [vp, ip, ct.c_char_p, ip]
    lib.fw_batch_get_shapes.argtypes = [vp, ct.POINTER(ip)]
    _LIB = lib
    return lib


def wavemaker_code(name: str) -> int:
    name = (name or "").strip()
    for k, v in WK.items():
        if name.startswith(k):
            return v
    if name == "" or any(name.startswith(k) for k in NO_SOURCE_WAVEMAKERS):
        return 0
    raise ValueError(f"wavemaker {name!r} has no per-step source on this path")


def config_from_state(st: dict, real_bytes: int) -> Config:
    """Build the scalar configuration from a state dict (reference names)."""
    c = Config()
    for n in INT_CFG:
        if n == "real_bytes":
            c.real_bytes = real_bytes
        elif n == "wavemaker":
            c.wavemaker = wavemaker_code(st.get("WaveMaker", ""))
        elif n in st:
            setattr(c, n, int(st[n]))
    for n in DBL_CFG:
        if n in st:
            setattr(c, n, float(st[n]))
    return c


class Solver:
    """One HBM-resident FUNWAVE-TVD domain on the current GPU."""

    def __init__(self, state: dict, precision: str = "sp", upload: bool = True, full_state: bool = False,
                 host_math: bool = False):
        self.lib = load_library()
        self.rb = 8 if precision == "dp" else 4
        self.rdt = np.float64 if self.rb == 8 else np.float32
        self.cfg = config_from_state(state, self.rb)
        self.ctx = ct.c_void_p()
        rc = self.lib.fw_create(ct.byref(self.cfg), ct.byref(self.ctx))
        if rc != FW_OK:
            raise FwError(rc, "fw_create failed (configuration rejected or device error)")
        self.M, self.N = self.cfg.Mloc, self.cfg.Nloc
        if full_state:
            self._check(self.lib.fw_set_option(self.ctx, FW_OPT_FULL_STATE, 1))
        if host_math:
            self._check(self.lib.fw_set_option(self.ctx, FW_OPT_HOST_MATH, 1))
        if upload:
            self.upload_state(state)

    # -- lifetime
    def close(self):
        if self.ctx:
            self.lib.fw_destroy(self.ctx)
            self.ctx = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != FW_OK:
            raise FwError(rc, self.lib.fw_strerror(self.ctx).decode())
        return rc

    # -- state transfer (numpy arrays indexed [j, i] = Fortran (I, J))
    def shape_of(self, name):
        if name in ("xmk_wk",):
            return (self.M,)
        if name in ("ymk_wk",):
            return (self.N,)
        if name == "OMGN2D":
            return (max(self.cfg.NumFreq, 1),)
        if name in ("Cm", "Sm"):
            return (self.cfg.NumFreq, self.N, self.M)
        fid = self.lib.fw_field_id(name.encode())
        if fid < 0:
            raise KeyError(name)
        if name in XFACE:
            return (self.N, self.M + 1)
        if name in YFACE:
            return (self.N + 1, self.M)
        return (self.N, self.M)

    def upload(self, name, arr):
        fid = self.lib.fw_field_id(name.encode())
        if fid < 0:
            raise KeyError(name)
        dt = np.int32 if name in INTF else self.rdt
        a = np.ascontiguousarray(np.asarray(arr, dtype=dt))
        if a.shape != self.shape_of(name):
            raise ValueError(f"{name}: shape {a.shape} != {self.shape_of(name)}")
        self._check(self.lib.fw_upload(self.ctx, fid, a.ctypes.data))

    def download(self, name):
        fid = self.lib.fw_field_id(name.encode())
        if fid < 0:
            raise KeyError(name)
        dt = np.int32 if name in INTF else self.rdt
        a = np.zeros(self.shape_of(name), dt)
        self._check(self.lib.fw_download(self.ctx, fid, a.ctypes.data))
        return a

    def upload_state(self, state: dict):
        for n in ALL_FIELDS:
            v = state.get(n)
            if isinstance(v, np.ndarray) and v.size > 0:
                if n in ("Cm", "Sm") and self.cfg.wavemaker != 2:
                    continue
                self.upload(n, v)
        self.set_clock(float(state.get("DT", 0.0)), float(state.get("TIME", 0.0)))

    def set_clock(self, dt, time):
        self._check(self.lib.fw_set_clock(self.ctx, dt, time))

    @property
    def clock(self):
        dt, t = ct.c_double(), ct.c_double()
        self.lib.fw_get_clock(self.ctx, ct.byref(dt), ct.byref(t))
        return dt.value, t.value

    # -- routines (reference names)
    def exchange(self):
        self._check(self.lib.fw_exchange(self.ctx))

    def estimate_dt(self):
        dt, t = ct.c_double(), ct.c_double()
        self._check(self.lib.fw_estimate_dt(self.ctx, ct.byref(dt), ct.byref(t)))
        return dt.value, t.value

    def save_stage0(self, which):
        self._check(self.lib.fw_save_stage0(self.ctx, which))

    def dispersion(self):
        self._check(self.lib.fw_cal_dispersion(self.ctx))

    def fluxes(self):
        self._check(self.lib.fw_fluxes(self.ctx))

    def sources(self):
        self._check(self.lib.fw_source_terms(self.ctx))

    def estimate_huv(self, istage):
        self._check(self.lib.fw_estimate_huv(self.ctx, istage))

    def update_mask(self):
        self._check(self.lib.fw_update_mask(self.ctx))

    def breaking(self):
        self._check(self.lib.fw_wave_breaking(self.ctx))

    def sponge(self):
        self._check(self.lib.fw_sponge_damping(self.ctx))

    def check_blowup(self):
        m = ct.c_double()
        rc = self.lib.fw_check_blowup(self.ctx, ct.byref(m))
        return rc, m.value

    def step(self, n=1):
        dt, t = ct.c_double(), ct.c_double()
        rc = self.lib.fw_step(self.ctx, n, ct.byref(dt), ct.byref(t))
        if rc == FW_E_BLOWUP:
            raise FwError(rc, self.lib.fw_strerror(self.ctx).decode())
        self._check(rc)
        return dt.value, t.value

    def stats(self):
        s = Stats()
        self._check(self.lib.fw_reduce_stats(self.ctx, ct.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def stats_noalloc(self):
        """fw_reduce_stats_noalloc: the same budget with no allocation"""
        s = Stats()
        self._check(self.lib.fw_reduce_stats_noalloc(self.ctx, ct.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def synchronize(self):
        self._check(self.lib.fw_synchronize(self.ctx))

    def device_bytes(self):
        b = ct.c_int64()
        self._check(self.lib.fw_device_bytes(self.ctx, ct.byref(b)))
        return b.value

    # -- measurement
    def profile(self, on=True):
        self._check(self.lib.fw_profile_enable(self.ctx, 1 if on else 0))

    def profile_reset(self):
        self._check(self.lib.fw_profile_reset(self.ctx))

    def profile_get(self, name):
        n, ms = ct.c_int64(), ct.c_double()
        rc = self.lib.fw_profile_get(self.ctx, name.encode(), ct.byref(n), ct.byref(ms))
        if rc == FW_E_NOTFOUND:
            return 0, 0.0
        self._check(rc)
        return n.value, ms.value

    def profile_names(self):
        buf = ct.create_string_buffer(4096)
        self._check(self.lib.fw_profile_names(self.ctx, buf, 4096))
        return [s for s in buf.value.decode().split(",") if s]

    def time_steps(self, n):
        ms = ct.c_double()
        self._check(self.lib.fw_time_steps(self.ctx, n, ct.byref(ms)))
        return ms.value


def _names(block):
    import re
    src = open(os.path.join(HERE, "csrc", "fw_fields.h")).read()
    body = src.split(f"#define {block}(X)")[1].split("\n#")[0]
    return re.findall(r"X\((\w+)\)", body)


CELL = _names("FW_CELL_FIELDS")
XFACE = _names("FW_XFACE_FIELDS")
YFACE = _names("FW_YFACE_FIELDS")
INTF = _names("FW_INT_FIELDS")
AUX = _names("FW_AUX_FIELDS")
ALL_FIELDS = CELL + XFACE + YFACE + INTF + AUX


class Batch:
    """Ensemble of Solver members advanced together (fw_batch_*): one launch
    per phase of the step over all members, one captured graph per step.
    Members need FIXED_DT; they stay usable (download, step, ...) and must
    outlive the batch."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (ct.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        self.b = ct.c_void_p()
        rc = self.lib.fw_batch_create(arr, len(self.members), ct.byref(self.b))
        if rc != FW_OK:
            raise FwError(rc, "fw_batch_create: " + self.members[0].lib.fw_strerror(self.members[0].ctx).decode())

    def _check(self, rc):
        if rc != FW_OK:
            raise FwError(rc, self.lib.fw_batch_strerror(self.b).decode())
        return rc

    def step(self, n=1):
        self._check(self.lib.fw_batch_step(self.b, n))

    def time_steps(self, n):
        ms = ct.c_double()
        self._check(self.lib.fw_batch_time_steps(self.b, n, ct.byref(ms)))
        return ms.value

    def status(self):
        st = (ct.c_int * len(self.members))()
        self._check(self.lib.fw_batch_status(self.b, st))
        return list(st)

    def set_graph(self, on=True):
        self._check(self.lib.fw_batch_set_graph(self.b, 1 if on else 0))

    def autotune(self, trial_steps=2):
        buf = ct.create_string_buffer(8192)
        self._check(self.lib.fw_batch_autotune(self.b, trial_steps, buf, 8192))
        return buf.value.decode()

    def shapes(self):
        s = (ct.c_int * 8)()
        self._check(self.lib.fw_batch_get_shapes(self.b, s))
        names = ("disp", "src_huv", "cell", "mask_brk")
        return {k: (s[2 * i], s[2 * i + 1]) for i, k in enumerate(names)}

    def close(self):
        if self.b:
            self.lib.fw_batch_destroy(self.b)
            self.b = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
