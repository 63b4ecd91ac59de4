"""Benchmark: FUNWAVE-TVD per-time-step solver on one MI355X.

Workload: the 2-D large-domain beach (2048 x 2048 cells, dx = dy = 2 m,
1:40 planar beach + Gaussian shoal + band-limited noise, irregular
directional JONSWAP waves through the WK_DATA2D internal wavemaker, sponges
on three sides, wetting-drying, dispersion and breaking on), synthetic
inputs built on the host by fwcases/fwsetup.  A step is one whole time step
(ESTIMATE_DT + three SSP-RK3 stages); the metric is interior
cells x steps / second / 1e6.

Prints one JSON line.  ``--gpus`` above 1 is refused: this build runs one
GPU per domain.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "funwave-tvd_amd"))

METRIC = "Mcell-steps/s at 2048^2; ensemble runs/hr learned vs plain at eta err<=2%"
HBM_PEAK = 8.0e12           # MI355X HBM3E, B/s
BF16_PEAK = 2.5e15          # MI355X dense bf16 MFMA, FLOP/s
B_ALG_WORDS = 93            # algorithmic words per interior cell-step (whole step)

# Algorithmic traffic of each kernel per unit of its launch: (real words
# read, int32 words read, real words written, unit), counting every distinct
# array element the routine must read or write once.  Units: "cell" =
# Mloc*Nloc, "xface" = (Mloc+1)*Nloc, "yface" = Mloc*(Nloc+1), "inner" =
# Mglob*Nglob.
KERNEL_TRAFFIC = {
    "flux_x": (8, 2, 3, "xface"),        # U V HU HV Eta U4 V4 DepthX, MASK MASK9 -> P Fx Gx
    "flux_y": (8, 2, 3, "yface"),
    "disp_deriv": (6, 1, 11, "cell"),    # U V Depth Eta P Q, MASK9 -> 8 second derivatives, ETAT ETAx ETAy
    "disp_terms": (17, 1, 6, "cell"),    # 8 second derivatives, ETAT ETAx ETAy, U V U0 V0 Depth Eta
                                         #   -> U4 V4 U1p V1p and the sums U1pp+U2+U3, V1pp+V2+V3
    "disp": (8, 1, 7, "cell"),           # one pass: U V U0 V0 Depth Eta P Q, MASK9 -> ETAT ETAx ETAy,
                                         #   U4 V4 and the two sums (the RK update forms the cross
                                         #   derivatives and U1p, V1p of its cell itself)
    "sources": (24, 2, 2, "inner"),
    "huv": (24, 2, 12, "cell"),
    # SourceTerms fused into ESTIMATE_HUV: U V U4 V4 Cd H P Q Eta DepthX DepthY,
    # the two dispersive sums, nu_break HU HV Eta0 Ubar0 Vbar0 Ubar Vbar Fx Fy
    # Gx Gy Depth, MASK MASK9 (the cross derivatives and U1p, V1p formed from
    # U, V, Depth) -> Eta Ubar Vbar H and the six line-solve coefficient arrays
    "src_huv": (26, 2, 10, "cell"),
    "trid": (6, 0, 2, "inner"),
    "uv_final": (3, 1, 4, "cell"),
    "mask": (4, 2, 2, "cell"),
    "brk1": (8, 0, 2, "cell"),
    "mask_mul": (6, 1, 5, "cell"),
}
# limiter of the kernels whose time is not set by HBM traffic (measured:
# VALU busy and achieved bytes in the committed counter profiles)
LIMITER = {"flux_x": "valu", "flux_y": "valu", "trid": "latency (serial recurrences)",
           "brk2": "latency (one workgroup)", "dt_final": "latency (one workgroup)",
           "max_final": "latency (one workgroup)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["sp", "dp"], default="sp")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--host-math", action="store_true",
                    help="Froude cap's ATAN2/COS/SIN by the host C library (FW_OPT_HOST_MATH)")
    ap.add_argument("--cpu-size", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--ensemble", action="store_true",
                    help="measure learned coarse vs plain fine ensemble runs/hour instead")
    ap.add_argument("--train-members", type=int, default=16)
    ap.add_argument("--eval-members", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--fine-steps", type=int, default=200)
    ap.add_argument("--workers", type=int, default=8,
                    help="ensemble members run at a time (host threads)")
    ap.add_argument("--batch", action="store_true",
                    help="batched ensemble instead: members of uneven size in one launch per phase")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--cpu-members-steps", type=int, default=3)
    return ap.parse_args()


def ensemble(args):
    """Learned coarse vs plain fine ensemble: conical-island solitary-wave
    members (INI_SOL); the coarse grid (dx = 4 m, Mglob ~ U[256,768], Nglob
    ~ U[128,512]) runs with the network correction after every step, the
    fine grid (dx = 2 m, twice the cells each way) is the plain method it is
    judged against and falls back to.  Each mode's members run together as
    one ensemble (fw_batch).  Prints one JSON line; the learned rate counts
    only if every member's delivered eta is within 2 % relative L2 of the
    2x2-averaged fine solution."""
    import numpy as np
    import fwlearn
    t0 = time.time()
    members = []
    for seed in range(5000, 5000 + args.train_members):
        c = fwlearn.member(seed)
        members.append(fwlearn.with_fixed_dt(c, fwlearn.stable_dt(c)))
    data = [fwlearn.make_samples(c, n_samples=16, every=6) for c in members]
    t_data = time.time() - t0
    params, losses = fwlearn.train(data, steps=args.train_steps, batch=16, size=64, lr=2e-3)
    t_train = time.time() - t0 - t_data
    evals = []
    for seed in range(6000, 6000 + args.eval_members):
        c = fwlearn.member(seed)
        evals.append(fwlearn.with_fixed_dt(c, fwlearn.stable_dt(c)))
    r = fwlearn.compare_batched(evals, args.fine_steps, params)
    rph = r["runs_per_hour"]
    ok = r["all_learned_within_bound"]
    out = {
        "metric": "ensemble runs/hr learned vs plain at eta err<=2%",
        "value": round(rph["learned"], 1) if ok else None, "unit": "runs/hour", "n_gpus": 1,
        "higher_is_better": True, "data": "synthetic", "dtype": "f32 solver, bf16 network",
        "accepted": ok, "plain_fine_runs_per_hour": round(rph["fine"], 1),
        "learned_runs_per_hour": round(rph["learned"], 1), "plain_coarse_runs_per_hour": round(rph["coarse"], 1),
        "speedup_learned_vs_fine": round(rph["learned"] / rph["fine"], 3),
        "fallbacks_to_fine": r["fallbacks"], "members_improved": r["members_improved"],
        "max_err_learned": round(r["max_err_learned"], 4), "max_err_coarse": round(r["max_err_coarse"], 4),
        "wall_s": {k: round(v, 3) for k, v in r["wall_s"].items()},
        "config": {"workload": "conical-island solitary-wave members (INI_SOL), fixed dt, coarse dx=4 m "
                               "(Mglob~U[256,768], Nglob~U[128,512]) vs fine dx=2 m, "
                               f"{args.fine_steps} fine steps ({args.fine_steps // 2} coarse) each; every mode's "
                               "members advanced together (fw_batch, one launch per phase), the learned runs "
                               "corrected after every step without the wet-cell mean of the eta correction "
                               "(moves no water); a learned member whose water volume drifts by more than "
                               f"{fwlearn.FALLBACK_DRIFT:g} or whose energy grows by more than "
                               f"x{fwlearn.FALLBACK_ENERGY:g} falls back to its fine run (counted in the "
                               "learned time)",
                   "train_members": args.train_members, "eval_members": args.eval_members,
                   "train_steps": args.train_steps, "loss_first": float(np.mean(losses[:20])),
                   "loss_last": float(np.mean(losses[-20:])),
                   "data_s": round(t_data, 1), "train_s": round(t_train, 1)},
        "members": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in m.items()} for m in r["members"]],
    }
    print(json.dumps(out))


def batch_states(n_members, precision):
    """The batched-ensemble workload: member m (seed 1000+m) with Mglob ~
    U[256,1024], Nglob ~ U[128,768], dx = 2 m, a planar slope ~ U[1:20,1:60]
    plus 1-4 Gaussian bumps, WK_REG waves of amplitude ~ U[0.2,1.0] m and
    period ~ U[6,14] s, a fixed time step of half the initial CFL step
    (rounded down to 1 ms)."""
    import numpy as np
    import fwcases
    import fwgpu
    import fwsetup
    cases, states = [], []
    for m in range(n_members):
        c = fwcases.member(m)
        st = fwsetup.build_state(c, precision)
        g = fwgpu.Solver(st, precision)
        dt, _ = g.estimate_dt()
        g.close()
        c.params["DT_fixed"] = float(np.floor(500.0 * dt) / 1000.0)
        cases.append(c)
        states.append(fwsetup.build_state(c, precision))
    return cases, states


def cpu_ensemble_baseline(cases, precision, steps, cores):
    """The compiled reference on the host cores: `cores` members at once,
    one serial process per core (each its own member of the same ensemble),
    `steps` timed steps each after one warm-up step; aggregate cell-steps
    per second over the concurrent members."""
    ref = os.path.join(REPO, "oracle", "_ref", f"fwref_{precision}")
    if not os.path.exists(ref):
        return None
    import shutil
    procs, dirs = [], []
    t0 = time.time()
    for c in cases[:cores]:
        d = tempfile.mkdtemp()
        c.write(d)
        dirs.append(d)
        env = dict(os.environ, FWREF_NSTEPS=str(1 + steps), FWREF_WARMUP="1", MALLOC_PERTURB_="255",
                   OMP_NUM_THREADS="1")
        procs.append(subprocess.Popen(["bash", "-c", f"ulimit -s unlimited; exec {ref} input.txt"], cwd=d, env=env,
                                      stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True))
    import re
    rate = 0.0
    per_member = []
    for c, p in zip(cases, procs):
        out, _ = p.communicate(timeout=900)
        line = [l for l in out.splitlines() if l.startswith("FWREF_TIMING")]
        if p.returncode != 0 or not line:
            return {"error": f"reference run failed rc={p.returncode}"}
        mt = re.search(r"steps=\s*(\d+)\s+seconds=\s*([0-9.Ee+-]+)", line[0])
        n, secs = int(mt.group(1)), float(mt.group(2))
        cells = c.params["Mglob"] * c.params["Nglob"]
        rate += cells * n / secs
        per_member.append(secs / n)
    for d in dirs:
        shutil.rmtree(d, ignore_errors=True)
    return {"value": round(rate / 1e6, 4), "unit": "Mcell-steps/s", "cores": len(procs), "kind": "reference",
            "sample": f"reference FUNWAVE-TVD ({'FP64' if precision == 'dp' else 'FP32'} serial build) on the first "
                      f"{len(procs)} members of the same ensemble, one process per core, all at once: {steps} steps "
                      f"each after 1 warm-up step ({time.time() - t0:.1f} s with set-up); mean "
                      f"{sum(per_member) / len(per_member):.3f} s per member-step"}


def batch_bench(args):
    """Batched ensemble: --members members of uneven size (batch_states)
    advanced --batch-steps whole steps through fw_batch (one launch per
    phase over all members, one graph per step, autotuned shapes), against
    the same members one after another with fw_step, and the reference on
    the host cores."""
    import fwgpu
    rb = 8 if args.precision == "dp" else 4
    t_setup = time.time()
    cases, states = batch_states(args.members, args.precision)
    gs = [fwgpu.Solver(st, args.precision) for st in states]
    cells = [g.cfg.Mglob * g.cfg.Nglob for g in gs]
    b = fwgpu.Batch(gs)
    report = b.autotune(trial_steps=2)
    t_setup = time.time() - t_setup
    b.step(args.warmup)
    sync()
    t0 = time.perf_counter()
    dev_ms = b.time_steps(args.batch_steps)
    sync()
    wall = time.perf_counter() - t0
    assert b.status() == [0] * len(gs), b.status()
    tot = sum(cells)
    mcs = tot * args.batch_steps / wall / 1e6
    # the same members one at a time (fw_step, one context each), a few steps
    k = 10
    t1 = time.perf_counter()
    for g in gs:
        g.step(k)
        g.synchronize()
    seq = time.perf_counter() - t1
    seq_mcs = tot * k / seq / 1e6
    step_s = wall / args.batch_steps
    alg = B_ALG_WORDS * rb * tot
    out = {
        "metric": "Mcell-steps/s, batched ensemble (all members)", "value": round(mcs, 3), "unit": "Mcell-steps/s",
        "n_gpus": 1, "steps": args.batch_steps, "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64" if rb == 8 else "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(gs)} ensemble members (seeds 1000..{999 + len(gs)}), Mglob~U[256,1024], "
                               "Nglob~U[128,768], dx=2 m, slope + bumps, WK_REG waves, west sponge, dispersion + "
                               "breaking, FIXED_DT; one launch per phase over all members, one hipGraph per step",
                   "members": len(gs), "cells_total": tot, "cells_min": min(cells), "cells_max": max(cells),
                   "parallelism": "all members on one GPU"},
        "roofline": {"bound": "hbm", "kernel": "whole step of every member",
                     "achieved": round(alg / step_s / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(alg / step_s / HBM_PEAK, 4), "traffic": None,
                     "alg_bytes_per_cell_step": B_ALG_WORDS * rb,
                     "device_ms_per_step": round(dev_ms / args.batch_steps, 4)},
        "ensemble_runs_per_hour": round(3600.0 / (step_s * args.batch_steps) * len(gs), 1),
        "members_one_at_a_time_Mcell_steps_s": round(seq_mcs, 3),
        "speedup_vs_one_at_a_time": round(mcs / seq_mcs, 3),
        "autotune": report.strip().splitlines()[-4:], "setup_s": round(t_setup, 1),
    }
    if not args.no_cpu_baseline:
        cores = min(16, os.cpu_count() or 1)
        cb = cpu_ensemble_baseline(cases, args.precision, args.cpu_members_steps, cores)
        if cb and "value" in cb:  # member runs (of the mean member size) per hour
            cb["ensemble_runs_per_hour"] = round(3600.0 * cb["value"] * 1e6 / (tot / len(gs) * args.batch_steps), 3)
        out["cpu_baseline"] = cb
    b.close()
    for g in gs:
        g.close()
    print(json.dumps(out))


def sync():
    """Device-wide barrier for work queued outside the solver's stream (only
    torch can have queued any, and only if something imported it)."""
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.synchronize()


# rocprofv3 names of the kernels behind each profiling name
ROCPROF_NAME = {"trid": "k_trid_lds<", "src_huv": "k_huv<{r}, true,", "huv": "k_huv<{r}, false,",
                "flux_x": "k_flux_x<", "flux_y": "k_flux_y", "disp_deriv": "k_disp_deriv<",
                "disp_terms": "k_disp_terms<", "disp": "k_disp_fused<", "sources": "k_sources<", "uv_final": "k_uv_final<",
                "mask": "k_mask<", "brk1": "k_brk1<", "mask_mul": "k_mask_mul<"}


def measured_traffic(kernel, rb, size):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, calibrated on a
    copy kernel), or None if there is none for this size / precision."""
    tag = f"beach{size}_{'fp64' if rb == 8 else 'fp32'}"
    path = os.path.join(REPO, "profiles", f"counters_{tag}.md")
    if kernel not in ROCPROF_NAME or not os.path.exists(path):
        return None, None
    want = ROCPROF_NAME[kernel].format(r="double" if rb == 8 else "float")
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0].startswith(want) and (("double" if rb == 8 else "float") in cells[0]):
            return float(cells[4]) * 1e6, os.path.relpath(path, REPO)
    return None, None


def cnn_roofline(n=1024, iters=20):
    """The learned correction's network (four bf16 MFMA convolutions) on the
    1024 x 1024 coarse grid of the 2048^2 workload: device time per forward
    pass against its roofline.  Each layer streams its input and output
    activations once (BYTE_PER_CELL: 8-channel bf16 input 16 B, three
    32-channel bf16 hidden activations written and read 64 + 64 B each, fp32
    output 12 B), so the pass does FLOP_PER_CELL / BYTE_PER_CELL ~ 100
    FLOP/B, below the bf16 ridge (2.5e15 / 8e12 ~ 312 FLOP/B): the bound is
    HBM; the MFMA rate is reported beside it."""
    import numpy as np
    import fwcnn
    net = fwcnn.Net(n, n, 1, fwcnn.init_params(0))
    net.forward(np.zeros((1, 5, n, n), np.float32))
    ms = net.time_forward(1, iters)
    net.close()
    cells = n * n
    flops = fwcnn.FLOP_PER_CELL * cells
    byts = fwcnn.BYTE_PER_CELL * cells
    s = ms * 1e-3
    return {"kernel": "k_conv x4 (bf16 MFMA 16x16x32, fp32 accumulate)", "cells": cells,
            "ms_per_forward": round(ms, 4), "bound": "hbm",
            "achieved": round(byts / s / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(byts / s / HBM_PEAK, 4), "byte_per_cell": fwcnn.BYTE_PER_CELL,
            "mfma_achieved_tflops": round(flops / s / 1e12, 2), "mfma_peak_tflops": BF16_PEAK / 1e12,
            "mfma_frac": round(flops / s / BF16_PEAK, 4), "flop_per_cell": fwcnn.FLOP_PER_CELL}


def kernel_traffic_bytes(name, rb, M, N, mg, ng):
    r, i, w, unit = KERNEL_TRAFFIC[name]
    units = {"cell": M * N, "xface": (M + 1) * N, "yface": M * (N + 1), "inner": mg * ng}[unit]
    return units * ((r + w) * rb + i * 4)


def _reference_runs(ref, case, steps, warm, cores):
    """Runs the compiled reference on `case` in `len(cores)` concurrent
    processes, process k pinned to core cores[k], each with its own input
    folder; returns [(steps, seconds)] from the reference's own timing
    lines, or None if a run failed."""
    import re
    dirs, procs = [], []
    try:
        for c in cores:
            td = tempfile.mkdtemp(prefix="fwref_")
            dirs.append(td)
            case.write(td)
            env = dict(os.environ, FWREF_NSTEPS=str(warm + steps), FWREF_WARMUP=str(warm),
                       MALLOC_PERTURB_="255", OMP_NUM_THREADS="1")
            procs.append(subprocess.Popen(["bash", "-c", f"ulimit -s unlimited; exec taskset -c {c} {ref} input.txt"],
                                          cwd=td, env=env, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                                          text=True))
        out = []
        for pr in procs:
            so, _ = pr.communicate(timeout=900)
            line = [ln for ln in so.splitlines() if ln.startswith("FWREF_TIMING")]
            if pr.returncode != 0 or not line:
                return None
            mt = re.search(r"steps=\s*(\d+)\s+seconds=\s*([0-9.Ee+-]+)", line[0])
            out.append((int(mt.group(1)), float(mt.group(2))))
        return out
    finally:
        for pr in procs:
            if pr.poll() is None:
                pr.kill()
        for td in dirs:
            shutil.rmtree(td, ignore_errors=True)


def cpu_baseline(args, precision):
    """The compiled reference (oracle/_ref) on a bounded sample of the same
    workload family, timed with its own monotonic clock: one process per
    host core of this GPU's share (up to 16), each a member of its own, all
    at once (the reference's serial build advances one domain per core), and
    one process alone."""
    ref = os.path.join(REPO, "oracle", "_ref", f"fwref_{precision}")
    if not os.path.exists(ref):
        return None
    import fwcases
    n = args.cpu_size
    case = fwcases.beach_2d(mglob=n, nglob=n, dx=2.0)
    warm = 2
    avail = sorted(os.sched_getaffinity(0))
    cores = avail[:min(16, len(avail))]
    t0 = time.time()
    one = _reference_runs(ref, case, args.cpu_steps, warm, cores[:1])
    many = _reference_runs(ref, case, args.cpu_steps, warm, cores) if len(cores) > 1 else one
    wall = time.time() - t0
    if not one or not many:
        return {"error": "reference run failed"}
    single = n * n * one[0][0] / one[0][1] / 1e6
    agg = sum(n * n * st / sec / 1e6 for st, sec in many)
    prec = "FP64" if precision == "dp" else "FP32"
    return {"value": round(agg, 3), "unit": "Mcell-steps/s", "cores": len(cores), "kind": "reference",
            "single_core": round(single, 3),
            "sample": f"reference FUNWAVE-TVD ({prec} serial build, amdflang -O2) on the same beach workload at "
                      f"{n}x{n} cells, dx=2 m, {args.cpu_steps} steps after {warm} warm-up steps: "
                      f"{len(cores)} processes at once, one per core (cores {cores[0]}-{cores[-1]} of "
                      f"{len(avail)} available), each advancing its own domain, summed; single_core: one "
                      f"process alone ({wall:.1f} s wall for both runs with set-up)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; this build runs one domain on one GPU "
              "(multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)
    if args.ensemble:
        ensemble(args)
        return
    if args.batch:
        batch_bench(args)
        return
    import numpy as np  # noqa: F401
    import fwcases
    import fwgpu
    import fwsetup

    t_setup = time.time()
    n = args.size
    case = fwcases.beach_2d(mglob=n, nglob=n, dx=2.0)
    st = fwsetup.build_state(case, args.precision)
    g = fwgpu.Solver(st, args.precision, host_math=args.host_math)
    t_setup = time.time() - t_setup
    rb = 8 if args.precision == "dp" else 4
    M, N = g.M, g.N

    g.step(args.warmup)
    sync()
    g.synchronize()
    t0 = time.perf_counter()
    dev_ms = g.time_steps(args.steps)
    g.synchronize()
    sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    mcs = n * n * args.steps / wall / 1e6

    # per-kernel device time (HIP events on the solver's stream)
    g.profile(True)
    g.profile_reset()
    g.step(args.profile_steps)
    kern = {}
    for name in g.profile_names():
        cnt, ms = g.profile_get(name)
        kern[name] = (cnt, ms)
    g.profile(False)
    step_s = wall / args.steps
    step_alg = B_ALG_WORDS * rb * n * n
    # Every kernel of the step, longest first: device ms per step, share of
    # the summed kernel time, its algorithmic bytes per launch over its
    # average launch time, and its limiter (from the committed counters).
    ksum = sum(v[1] for v in kern.values())
    breakdown = []
    traffic_step, traffic_src = 0.0, None
    for name, (cnt, ms) in sorted(kern.items(), key=lambda kv: -kv[1][1]):
        row = {"kernel": name, "ms_per_step": round(ms / args.profile_steps, 4),
               "launches_per_step": round(cnt / args.profile_steps, 2), "share": round(ms / ksum, 4),
               "bound": LIMITER.get(name, "hbm")}
        if name in KERNEL_TRAFFIC:
            alg = kernel_traffic_bytes(name, rb, M, N, n, n)
            row["alg_GBps"] = round(alg / (ms / cnt / 1e3) / 1e9, 1)
            row["alg_frac_of_hbm_peak"] = round(alg / (ms / cnt / 1e3) / HBM_PEAK, 4)
        t, src = measured_traffic(name, rb, n)
        if t is not None:
            row["counter_MB_per_launch"] = round(t / 1e6, 1)
            traffic_step += t * cnt / args.profile_steps
            traffic_src = src
        breakdown.append(row)
    out = {
        "metric": METRIC, "value": round(mcs, 3), "unit": "Mcell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if rb == 8 else "f32", "data": "synthetic",
        "config": {"workload": f"2-D beach + Gaussian shoal + noise, JONSWAP 200-component WK_DATA2D "
                               f"waves, {n}x{n} cells dx=2 m, sponges W/S/N, wetting-drying, "
                               f"dispersion + breaking, whole RK3 steps",
                   "Mglob": n, "Nglob": n, "precision": "FP64" if rb == 8 else "FP32",
                   "line_solver": "thomas sweep, c' and d' recurrences on two waves of one workgroup "
                                  "(LDS-staged; bitwise the reference's recurrence)",
                   "froude_cap_math": "host C library" if args.host_math else "device math library",
                   "parallelism": "one domain on one GPU"},
        # The whole step against its canonical algorithmic traffic:
        # 93 words per interior cell-step (372 B FP32, 744 B FP64), i.e.
        # the fields that must cross HBM when every stage-local intermediate
        # stays on chip; traffic = the rocprofv3 counter bytes of one step
        # (every kernel's calibrated FETCH_SIZE + WRITE_SIZE per launch x its
        # launches per step).
        "roofline": {"bound": "hbm", "kernel": "whole step (every kernel of ESTIMATE_DT + 3 RK stages)",
                     "achieved": round(step_alg / step_s / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(step_alg / step_s / HBM_PEAK, 4),
                     "traffic": round(traffic_step) if traffic_src else None,
                     "traffic_source": traffic_src,
                     "alg_bytes_per_step": step_alg, "alg_bytes_per_cell_step": B_ALG_WORDS * rb,
                     "device_ms_per_step": round(dev_ms / args.steps, 4)},
        "kernels": breakdown,
        "setup_s": round(t_setup, 1),
    }
    out["learned_correction"] = cnn_roofline()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, args.precision)
    g.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
