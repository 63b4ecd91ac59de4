"""Python binding of the learned coarse-grid correction (include/fwcnn.h).

The network is 5 -> 32 -> 32 -> 32 -> 3 channels of 3x3 convolutions with
zero padding and ReLU between layers, evaluated on the gfx950 matrix cores
(bf16 operands, fp32 accumulation).  Inputs are the coarse state features
(eta, u, v, depth, mask) scaled by FEATURE_SCALE; outputs are the one-step
corrections (d_eta, d_u, d_v) divided by OUTPUT_SCALE.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

import fwgpu

NPARAMS = 9 * 5 * 32 + 32 + 2 * (9 * 32 * 32 + 32) + 9 * 32 * 3 + 3
SHAPES = [("W1", (9, 5, 32)), ("b1", (32,)), ("W2", (9, 32, 32)), ("b2", (32,)),
          ("W3", (9, 32, 32)), ("b3", (32,)), ("W4", (9, 32, 3)), ("b4", (3,))]
EXPORTS = ["fw_cnn_create", "fw_cnn_destroy", "fw_cnn_strerror", "fw_cnn_set_params",
           "fw_cnn_get_params", "fw_cnn_forward", "fw_cnn_train_step", "fw_cnn_correct", "fw_cnn_set_conserve",
           "fw_cnn_time_forward"]
# multiply-adds x 2 per cell of the four layers (algorithmic, unpadded channels)
FLOP_PER_CELL = 2 * 9 * (5 * 32 + 32 * 32 + 32 * 32 + 32 * 3)
# HBM bytes per cell of a forward pass, each activation written and read once:
# 8-channel bf16 input 16, three 32-channel bf16 activations 3 x (64 + 64),
# fp32 output 3 x 4
BYTE_PER_CELL = 16 + 3 * (64 + 64) + 12
# feature scales: eta / 1 m, u, v / 1 m/s, depth / 10 m, mask
FEATURE_SCALE = np.array([1.0, 1.0, 1.0, 0.1], np.float32)
# outputs are corrections in units of 1 mm (eta) and 1 mm/s (u, v)
OUTPUT_SCALE = np.array([1e-3, 1e-3, 1e-3], np.float32)


def unflatten(p: np.ndarray) -> dict:
    out, k = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[k:k + n].reshape(shp)
        k += n
    assert k == NPARAMS
    return out


def flatten(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[n], np.float32).ravel() for n, _ in SHAPES]).astype(np.float32)


def init_params(seed: int = 0, last_scale: float = 0.1) -> np.ndarray:
    """He-normal weights, zero biases; the last layer scaled down so that an
    untrained network makes small corrections."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan_in = shp[0] * shp[1]
            w = rng.standard_normal(shp) * np.sqrt(2.0 / fan_in)
            d[name] = w * (last_scale if name == "W4" else 1.0)
        else:
            d[name] = np.zeros(shp)
    return flatten(d)


def _lib():
    lib = fwgpu.load_library()
    if not getattr(lib, "_fwcnn_typed", False):
        vp, fp, ip = ct.c_void_p, ct.POINTER(ct.c_float), ct.c_int
        lib.fw_cnn_create.argtypes = [ip, ip, ip, ct.POINTER(vp)]
        lib.fw_cnn_destroy.argtypes = [vp]
        lib.fw_cnn_strerror.argtypes = [vp]
        lib.fw_cnn_strerror.restype = ct.c_char_p
        lib.fw_cnn_set_params.argtypes = [vp, fp, ip]
        lib.fw_cnn_get_params.argtypes = [vp, fp, ip]
        lib.fw_cnn_forward.argtypes = [vp, fp, fp, ip]
        lib.fw_cnn_train_step.argtypes = [vp, fp, fp, ip, ct.c_float, fp, fp]
        lib.fw_cnn_correct.argtypes = [vp, vp, fp, fp]
        lib.fw_cnn_set_conserve.argtypes = [vp, ip]
        lib.fw_cnn_time_forward.argtypes = [vp, ip, ip, ct.POINTER(ct.c_double)]
        lib._fwcnn_typed = True
    return lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


class Net:
    """Network buffers for batches of up to max_batch H x W samples."""

    def __init__(self, H: int, W: int, max_batch: int = 1, params: np.ndarray | None = None):
        self.lib = _lib()
        self.H, self.W, self.max_batch = H, W, max_batch
        self.ptr = ct.c_void_p()
        rc = self.lib.fw_cnn_create(H, W, max_batch, ct.byref(self.ptr))
        if rc != fwgpu.FW_OK:
            raise fwgpu.FwError(rc, "fw_cnn_create failed")
        self.set_params(init_params() if params is None else params)

    def _check(self, rc):
        if rc != fwgpu.FW_OK:
            raise fwgpu.FwError(rc, self.lib.fw_cnn_strerror(self.ptr).decode())

    def close(self):
        if self.ptr:
            self.lib.fw_cnn_destroy(self.ptr)
            self.ptr = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self.lib.fw_cnn_set_params(self.ptr, _fptr(p), NPARAMS))

    def params(self) -> np.ndarray:
        p = np.empty(NPARAMS, np.float32)
        self._check(self.lib.fw_cnn_get_params(self.ptr, _fptr(p), NPARAMS))
        return p

    def forward(self, x: np.ndarray) -> np.ndarray:
        """x: [B, 5, H, W] -> [B, 3, H, W]"""
        x = np.ascontiguousarray(x, np.float32)
        B = x.shape[0]
        assert x.shape == (B, 5, self.H, self.W)
        out = np.empty((B, 3, self.H, self.W), np.float32)
        self._check(self.lib.fw_cnn_forward(self.ptr, _fptr(x), _fptr(out), B))
        return out

    def train_step(self, x: np.ndarray, target: np.ndarray, lr: float, want_grad: bool = False):
        """One Adam step on the mean squared error; returns (loss, grad|None)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        B = x.shape[0]
        loss = ct.c_float()
        g = np.empty(NPARAMS, np.float32) if want_grad else None
        self._check(self.lib.fw_cnn_train_step(self.ptr, _fptr(x), _fptr(t), B, lr, ct.byref(loss),
                                               _fptr(g) if want_grad else None))
        return float(loss.value), g

    def time_forward(self, B: int = 1, iters: int = 20) -> float:
        """Device milliseconds per forward pass over B resident samples."""
        ms = ct.c_double()
        self._check(self.lib.fw_cnn_time_forward(self.ptr, B, iters, ct.byref(ms)))
        return ms.value

    def set_conserve(self, on: bool = True):
        """Take the wet-cell mean of the eta correction out of it (no net
        change of water volume)."""
        self._check(self.lib.fw_cnn_set_conserve(self.ptr, 1 if on else 0))

    def correct(self, solver: "fwgpu.Solver", fs=FEATURE_SCALE, os_=OUTPUT_SCALE):
        """Queue the in-place correction of a coarse solver's state."""
        fs = np.ascontiguousarray(fs, np.float32)
        os_ = np.ascontiguousarray(os_, np.float32)
        self._check(self.lib.fw_cnn_correct(self.ptr, solver.ctx, _fptr(fs), _fptr(os_)))
