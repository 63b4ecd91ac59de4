"""Synthetic FUNWAVE-TVD workloads written in the reference's own input format.

Each builder returns a ``Case``: the text of an ``input.txt`` (key = value
lines read by the reference READ_INPUT, src/io.F:149-3773 via
src/mod_input.F:271-308) plus the data files it names (ASCII depth grid,
``Mglob`` values per row and ``Nglob`` rows as read at src/init.F:704-710;
2-D spectrum file in the WK_DATA2D layout read at src/wavemaker.F:197-215).
``Case.write(dir)`` puts them on disk so that the reference solver and this
package read identical inputs.

Workloads (sizes are parameters so tests can shrink them):
  beach_1d     quasi-1D planar beach, WK_REG internal wavemaker, periodic y
  beach_2d     2-D beach + Gaussian shoal + band-limited noise, JONSWAP
               directional spectrum through WK_DATA2D, sponges on 3 sides
  member       one ensemble member: slope + 1-4 Gaussian bumps, WK_REG
  island       solitary wave (INI_SOL) running up a conical island
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Case:
    name: str
    params: dict
    depth: np.ndarray | None = None          # (Nglob, Mglob), row j = y index
    spectrum: dict | None = None             # WK_DATA2D content
    files: dict = field(default_factory=dict)

    @property
    def mglob(self) -> int:
        return int(self.params["Mglob"])

    @property
    def nglob(self) -> int:
        return int(self.params["Nglob"])

    def input_text(self) -> str:
        lines = [f"TITLE = {self.name}"]
        for k, v in self.params.items():
            if isinstance(v, bool):
                v = "T" if v else "F"
            elif isinstance(v, float):
                v = repr(v)
            lines.append(f"{k} = {v}")
        return "\n".join(lines) + "\n"

    def write(self, path: str) -> str:
        os.makedirs(path, exist_ok=True)
        if self.depth is not None:
            with open(os.path.join(path, "depth.txt"), "w") as f:
                # 17 significant digits: the text round-trips the double exactly
                np.savetxt(f, np.asarray(self.depth, np.float64), fmt="%.16e")
        if self.spectrum is not None:
            s = self.spectrum
            with open(os.path.join(path, "spectrum.txt"), "w") as f:
                nf, nd = len(s["freq"]), len(s["dire"])
                f.write(f"{nf} {nd}\n{s['peak_period']:.9e}\n")
                for x in s["freq"]:
                    f.write(f"{x:.9e}\n")
                for x in s["dire"]:
                    f.write(f"{x:.9e}\n")
                for i in range(nd):
                    f.write(" ".join(f"{a:.9e}" for a in s["amp"][:, i]) + "\n")
                for i in range(nd):
                    f.write(" ".join(f"{a:.9e}" for a in s["phase_deg"][:, i]) + "\n")
        p = os.path.join(path, "input.txt")
        with open(p, "w") as f:
            f.write(self.input_text())
        return p


def _common(mglob, nglob, dx, dy, total_time, out="output/"):
    return {
        "PX": 1, "PY": 1,
        "Mglob": int(mglob), "Nglob": int(nglob),
        "DX": float(dx), "DY": float(dy),
        "TOTAL_TIME": float(total_time),
        "PLOT_INTV": 1.0e9, "SCREEN_INTV": 1.0e9, "PLOT_INTV_STATION": 1.0e9,
        "RESULT_FOLDER": out,
        "DISPERSION": True,
        "CFL": 0.5, "FroudeCap": 3.0, "MinDepth": 0.01,
        "VISCOSITY_BREAKING": True, "Cbrk1": 0.65, "Cbrk2": 0.35,
        "Cd": 0.0,
        "ETA": True, "U": True, "V": True, "MASK": True,
    }


def beach_1d(mglob=1024, nglob=4, total_time=60.0, xslp=400.0, xc_wk=200.0,
             amp=0.5, period=8.0, sponge_west=100.0):
    """Quasi-1D planar beach driven by the WK_REG source (config of the
    reference's simple_cases/surface_wave_1d/input_files/input_reg.txt):
    10 m flat bottom, 1:30 slope from ``xslp`` (src/init.F:750-776)."""
    p = _common(mglob, nglob, 1.0, 1.0, total_time)
    p.update({
        "DEPTH_TYPE": "SLOPE", "DEPTH_FLAT": 10.0, "SLP": 1.0 / 30.0,
        "Xslp": float(xslp),
        "WAVEMAKER": "WK_REG", "DEP_WK": 10.0, "Xc_WK": float(xc_wk),
        "Yc_WK": 0.0, "Tperiod": float(period), "AMP_WK": float(amp),
        "Theta_WK": 0.0, "Delta_WK": 3.0, "Time_ramp": 1.0,
        "PERIODIC": True,
        "DIRECT_SPONGE": True, "FRICTION_SPONGE": False,
        "DIFFUSION_SPONGE": False,
        "Sponge_west_width": float(sponge_west), "Sponge_east_width": 0.0,
        "Sponge_south_width": 0.0, "Sponge_north_width": 0.0,
        "R_sponge": 0.85, "A_sponge": 5.0,
    })
    return Case("beach_1d", p)


def _gauss_filter_periodic(white, sigma_cells):
    """Gaussian low-pass through the FFT (periodic wrap)."""
    ny, nx = white.shape
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    g = np.exp(-2.0 * (np.pi ** 2) * (sigma_cells ** 2) * (kx ** 2 + ky ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(white) * g))


def jonswap_2d(hs=1.0, tp=10.0, gamma=3.3, sigma_theta_deg=20.0, nfreq=20,
               ndir=10, seed=2):
    """Discrete directional JONSWAP spectrum: amplitudes per (freq, dir) bin
    with sum(a^2)/2 = (Hs/4)^2 and a cos^2s spreading of width sigma_theta;
    phases uniform in [0, 360) degrees from ``seed``."""
    fp = 1.0 / tp
    f = np.linspace(0.6 * fp, 2.4 * fp, nfreq)
    df = f[1] - f[0]
    sig = np.where(f <= fp, 0.07, 0.09)
    r = np.exp(-((f - fp) ** 2) / (2.0 * sig ** 2 * fp ** 2))
    s_f = (f ** -5) * np.exp(-1.25 * (fp / f) ** 4) * gamma ** r
    th = np.linspace(-45.0, 45.0, ndir)
    dth = th[1] - th[0] if ndir > 1 else 1.0
    sth = np.deg2rad(sigma_theta_deg)
    s_par = max(2.0 / sth ** 2 - 1.0, 1.0)
    d = np.cos(np.deg2rad(th) / 2.0) ** (2.0 * s_par)
    d = d / (d.sum() * dth)
    e = s_f[:, None] * d[None, :] * df * dth
    m0 = (hs / 4.0) ** 2
    e *= m0 / e.sum()
    amp = np.sqrt(2.0 * e)
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0.0, 360.0, size=amp.shape)
    return {"freq": f, "dire": th, "amp": amp, "phase_deg": phase,
            "peak_period": tp}


def beach_2d(mglob=2048, nglob=2048, dx=2.0, seed=1, phase_seed=2,
             total_time=1.0e9, nfreq=20, ndir=10, flat_depth=20.0,
             x_toe=None, slope=1.0 / 40.0, shoal=(2400.0, None, 3.0, 150.0),
             noise_std=0.2, noise_corr=50.0, xc_wk=600.0, sponge=300.0):
    """2-D planar beach + Gaussian shoal + band-limited noise, irregular
    directional waves from a JONSWAP spectrum fed through WK_DATA2D.

    At the default sizes this is the 2048 x 2048, dx=2 m large-domain
    workload; smaller ``mglob/nglob`` shrink it for tests (x positions scale
    with the domain length)."""
    lx = mglob * dx
    ly = nglob * dx
    scale = lx / 4096.0
    if x_toe is None:
        x_toe = 3000.0 * scale
    x = (np.arange(mglob) + 0.5) * dx
    y = (np.arange(nglob) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    depth = np.full((nglob, mglob), flat_depth)
    depth = np.where(X > x_toe, flat_depth - slope * (X - x_toe), depth)
    sx, sy, sh, ss = shoal
    sx = sx * scale
    sy = ly / 2.0 if sy is None else sy
    ss = ss * scale
    depth = depth - sh * np.exp(-((X - sx) ** 2 + (Y - sy) ** 2) / (2 * ss ** 2))
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((nglob, mglob))
    filt = _gauss_filter_periodic(white, max(noise_corr * scale / dx, 0.5))
    filt = filt / max(filt.std(), 1e-30) * noise_std
    depth = depth + filt
    spec = jonswap_2d(nfreq=nfreq, ndir=ndir, seed=phase_seed)
    p = _common(mglob, nglob, dx, dx, total_time)
    sp = sponge * scale
    p.update({
        "DEPTH_TYPE": "DATA", "DEPTH_FILE": "depth.txt",
        "WAVEMAKER": "WK_DATA2D", "WaveCompFile": "spectrum.txt",
        "DEP_WK": float(flat_depth), "Xc_WK": float(xc_wk * scale),
        "Yc_WK": 0.0, "Delta_WK": 3.0, "Time_ramp": 1.0,
        "PERIODIC": False,
        "DIRECT_SPONGE": True, "FRICTION_SPONGE": False,
        "DIFFUSION_SPONGE": False,
        "Sponge_west_width": float(sp), "Sponge_east_width": 0.0,
        "Sponge_south_width": float(sp), "Sponge_north_width": float(sp),
        "R_sponge": 0.85, "A_sponge": 5.0,
    })
    return Case("beach_2d", p, depth=depth, spectrum=spec)


def member(seed, dx=2.0, mrange=(256, 1024), nrange=(128, 768),
           total_time=1.0e9):
    """One ensemble member: planar slope ~U[1:60,1:20] plus 1-4 Gaussian
    bumps, WK_REG waves with amplitude ~U[0.2,1.0] m, period ~U[6,14] s."""
    rng = np.random.default_rng(1000 + seed)
    mglob = int(rng.integers(mrange[0], mrange[1] + 1))
    nglob = int(rng.integers(nrange[0], nrange[1] + 1))
    slope = 1.0 / rng.uniform(20.0, 60.0)
    amp = float(rng.uniform(0.2, 1.0))
    per = float(rng.uniform(6.0, 14.0))
    lx, ly = mglob * dx, nglob * dx
    x = (np.arange(mglob) + 0.5) * dx
    y = (np.arange(nglob) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    flat = 12.0
    x_toe = 0.35 * lx
    depth = np.where(X > x_toe, flat - slope * (X - x_toe), flat)
    for _ in range(int(rng.integers(1, 5))):
        bx, by = rng.uniform(0.3 * lx, 0.8 * lx), rng.uniform(0, ly)
        bh, bs = rng.uniform(0.5, 3.0), rng.uniform(10.0, 60.0)
        depth = depth - bh * np.exp(-((X - bx) ** 2 + (Y - by) ** 2) / (2 * bs ** 2))
    p = _common(mglob, nglob, dx, dx, total_time)
    p.update({
        "DEPTH_TYPE": "DATA", "DEPTH_FILE": "depth.txt",
        "WAVEMAKER": "WK_REG", "DEP_WK": flat, "Xc_WK": float(0.15 * lx),
        "Yc_WK": 0.0, "Tperiod": per, "AMP_WK": amp, "Theta_WK": 0.0,
        "Delta_WK": 3.0, "Time_ramp": 1.0,
        "PERIODIC": False,
        "DIRECT_SPONGE": True, "Sponge_west_width": float(0.08 * lx),
        "Sponge_east_width": 0.0, "Sponge_south_width": 0.0,
        "Sponge_north_width": 0.0, "R_sponge": 0.85, "A_sponge": 5.0,
    })
    return Case(f"member_{seed}", p, depth=depth)


def island(seed, dx=4.0, mrange=(256, 768), nrange=(128, 512),
           total_time=1.0e9, mglob=None, nglob=None, r_base=None, h0=None):
    """Solitary wave (INI_SOL, src/init.F:907-910) running up a seeded
    conical island (slope 1:4, base radius ~U[50,200] m, 5 % noise)."""
    rng = np.random.default_rng(9000 + seed)
    if mglob is None:
        mglob = int(rng.integers(mrange[0], mrange[1] + 1))
    if nglob is None:
        nglob = int(rng.integers(nrange[0], nrange[1] + 1))
    h0_draw = float(rng.uniform(0.32, 10.0))
    ah = float(rng.uniform(0.05, 0.3))
    r_draw = float(rng.uniform(50.0, 200.0))
    h0 = h0_draw if h0 is None else float(h0)
    r_base = r_draw if r_base is None else float(r_base)
    lx, ly = mglob * dx, nglob * dx
    x = (np.arange(mglob) + 0.5) * dx
    y = (np.arange(nglob) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    cx, cy = 0.6 * lx, 0.5 * ly
    r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2)
    top = h0 + 0.5 * h0  # the cone rises above still water by h0/2
    cone = np.clip((r_base - r) / 4.0, 0.0, top)
    cone = cone * (1.0 + 0.05 * rng.standard_normal(cone.shape))
    depth = h0 - cone
    p = _common(mglob, nglob, dx, dx, total_time)
    p.update({
        "DEPTH_TYPE": "DATA", "DEPTH_FILE": "depth.txt",
        "WAVEMAKER": "INI_SOL", "AMP": ah * h0, "DEP": h0,
        "XWAVEMAKER": float(0.25 * lx),
        "PERIODIC": False, "DIRECT_SPONGE": False,
    })
    return Case(f"island_{seed}", p, depth=depth)
