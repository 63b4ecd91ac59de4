"""Learned coarse-grid ensembles: training data, training and the
runs-per-hour comparison of learned coarse runs against plain fine runs.

A member is a FUNWAVE-TVD case run at its own spacing dx ("fine") and at
2*dx ("coarse", the 2x2 block mean of the fine bathymetry), both with a
fixed time step, the coarse one twice the fine one, so that one coarse step
spans exactly two fine steps.  The network (fwcnn) predicts, from the coarse
state right after a coarse step, the difference between the 2x2-restricted
fine state and that coarse state; the learned run adds it after every coarse
step (fw_cnn_correct).  A learned member counts only if its final surface
elevation is within ERR_BOUND relative L2 of the restricted fine solution.
Surface elevations are compared as water surfaces over wet cells (dry cells
hold eta = -MinDepth - Depth in the reference, which would otherwise swamp
the norm with land); the fine one is the wet-cell mean of each 2x2 block.
"""
from __future__ import annotations

import copy
import time

import numpy as np

import fwcases
import fwcnn
import fwgpu
import fwsetup

ERR_BOUND = 0.02
STATE_IN = ("Eta", "U", "V")


def interior(a):
    return a[3:-3, 3:-3]


def restrict(a):
    """2x2 block mean of an [N, M] interior array (N, M even)."""
    n, m = a.shape
    return a.reshape(n // 2, 2, m // 2, 2).mean(axis=(1, 3))


def restrict_wet(a, mask):
    """2x2 block mean over the wet cells of each block (0 where the whole
    block is dry) and the block's any-wet mask.  Dry cells hold the land
    height (eta = -MinDepth - Depth, Depth < 0 on land), which a plain mean
    would mix into the water surface."""
    w = np.asarray(mask, np.float64)
    s = restrict(np.asarray(a, np.float64) * w)
    c = restrict(w)
    return np.where(c > 0, s / np.where(c > 0, c, 1.0), 0.0), (c > 0).astype(np.float64)


def restricted_fields(get):
    """Eta, U, V, Ubar, Vbar of a fine state (get(name) -> interior array),
    restricted for a coarse state: wet-cell means; a block without wet cells
    takes the plain mean of eta (land height) and zero velocities."""
    m = get("MASK")
    out = []
    for k in STATE_IN + ("Ubar", "Vbar"):
        a, wet = restrict_wet(get(k), m)
        if k == "Eta":
            a = np.where(wet > 0, a, restrict(get(k)))
        out.append(a)
    return out


def with_fixed_dt(case, dt):
    c = copy.deepcopy(case)
    c.params["DT_fixed"] = float(dt)
    return c


def coarse_case(fine):
    """The same member on a grid twice as coarse: 2x2 block-mean depth,
    doubled spacing and fixed time step."""
    c = copy.deepcopy(fine)
    p = c.params
    if p["Mglob"] % 2 or p["Nglob"] % 2:
        raise ValueError("fine grid must have even Mglob, Nglob")
    p["Mglob"] //= 2
    p["Nglob"] //= 2
    p["DX"] = 2.0 * p["DX"]
    p["DY"] = 2.0 * p["DY"]
    if "DT_fixed" in p:
        p["DT_fixed"] = 2.0 * p["DT_fixed"]
    c.depth = restrict(np.asarray(fine.depth, np.float64))
    c.name = fine.name + "_coarse"
    return c


# The learned ensemble's members: the coarse grid has spacing 4 m and
# Mglob ~ U[256, 768], Nglob ~ U[128, 512] cells; the fine grid (the plain
# method the learned runs are judged against, and fall back to) has half
# the spacing and twice the cells in each direction.
COARSE_DX = 4.0
COARSE_M = (256, 768)
COARSE_N = (128, 512)


def member(seed, dx=COARSE_DX / 2, mrange=None, nrange=None, dt=None, **kw):
    """A seeded conical-island member (INI_SOL solitary wave) on the fine
    grid: spacing dx, sizes twice the coarse draws (even, so that the 2 x 2
    restriction is exact), and, if dt is given, a fixed time step."""
    rng = np.random.default_rng(70000 + seed)
    mr = mrange or (2 * COARSE_M[0], 2 * COARSE_M[1])
    nr = nrange or (2 * COARSE_N[0], 2 * COARSE_N[1])
    m = kw.pop("mglob", None) or 2 * int(rng.integers(mr[0] // 2, mr[1] // 2 + 1))
    n = kw.pop("nglob", None) or 2 * int(rng.integers(nr[0] // 2, nr[1] // 2 + 1))
    c = fwcases.island(seed, dx=dx, mglob=m, nglob=n, **kw)
    return with_fixed_dt(c, dt) if dt else c


def stable_dt(case, cfl=0.5):
    """A fixed step below the CFL limit of the initial state (sqrt(g h_max)
    plus the solitary wave's particle speed), rounded down to 1 ms."""
    h = float(np.max(case.depth)) + float(case.params.get("AMP", 0.0))
    c = np.sqrt(9.81 * h) * 1.3
    return float(np.floor(1000.0 * cfl * 0.8 * case.params["DX"] / c) / 1000.0)


def set_state(solver, base, eta, u, v, ubar=None, vbar=None):
    """Overwrite the interior surface and velocities (and, if given, the
    depth-integrated Ubar, Vbar) of a solver built from `base` (a host state
    of fwsetup) and refresh what INITIALIZATION derives from them
    (init.F:1000-1026, 1158-1163): wet/dry mask, dry-cell surface, H, HU,
    HV, MASK9, and Ubar = HU, Vbar = HV when not given.  Ghost cells copy the
    nearest interior cell; the solver's EXCHANGE refills them at the next
    step."""
    rdt = np.float64 if base["SP_BYTES"] == 8 else np.float32
    dep = base["Depth"].astype(np.float64)
    N, M = dep.shape

    def full(a):
        return np.pad(np.asarray(a, np.float64), 3, mode="edge")

    e, uu, vv = full(eta), full(u), full(v)
    mindep = float(base["MinDepth"])
    mask = np.where(e < -dep, 0, 1).astype(np.int32)
    e = np.where(mask == 0, -mindep - dep, e)
    uu = uu * mask
    vv = vv * mask
    h = np.maximum(e * float(base["Gamma3"]) + dep, float(base["MinDepthFrc"]))
    m9 = np.zeros((N, M), np.int32)
    prod = np.ones((N - 2, M - 2), np.int32)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            prod = prod * mask[1 + dj:N - 1 + dj, 1 + di:M - 1 + di]
    m9[3:-3, 3:-3] = prod[2:-2, 2:-2]
    ub = h * uu if ubar is None else full(ubar) * mask
    vb = h * vv if vbar is None else full(vbar) * mask
    fields = {"Eta": e, "U": uu, "V": vv, "H": h, "HU": h * uu, "HV": h * vv, "Ubar": ub, "Vbar": vb}
    for k, a in fields.items():
        solver.upload(k, a.astype(rdt))
    solver.upload("MASK", mask)
    solver.upload("MASK9", m9)


def features(solver, base) -> np.ndarray:
    """[5, H, W] network input of a coarse solver's current state."""
    fs = fwcnn.FEATURE_SCALE
    e, u, v = (interior(solver.download(k)).astype(np.float32) for k in STATE_IN)
    d = interior(base["Depth"]).astype(np.float32)
    m = interior(solver.download("MASK")).astype(np.float32)
    return np.stack([e * fs[0], u * fs[1], v * fs[2], d * fs[3], m]).astype(np.float32)


def make_samples(fine_case, n_samples, every=4, precision="sp"):
    """(inputs [S, 5, H, W], targets [S, 3, H, W]) along one fine trajectory:
    every `every` coarse steps the coarse solver is set to the restricted
    fine state, takes one step, and the target is the restricted fine state
    two fine steps later minus that coarse prediction (in output units)."""
    cc = coarse_case(fine_case)
    bf, bc = fwsetup.build_state(fine_case, precision), fwsetup.build_state(cc, precision)
    fine, coarse = fwgpu.Solver(bf, precision), fwgpu.Solver(bc, precision)
    xs, ys = [], []
    osc = fwcnn.OUTPUT_SCALE
    for _ in range(n_samples):
        if every > 1:
            fine.step(2 * (every - 1))
        now = restricted_fields(lambda k: interior(fine.download(k)).astype(np.float64))
        set_state(coarse, bc, *now)
        coarse.set_clock(*fine.clock)
        coarse.step(1)
        fine.step(2)
        wf = interior(fine.download("MASK")).astype(np.float64)
        nxt = [restrict_wet(interior(fine.download(k)).astype(np.float64), wf)[0] for k in STATE_IN]
        wet = interior(coarse.download("MASK")).astype(np.float64)
        pred = [interior(coarse.download(k)).astype(np.float64) * wet for k in STATE_IN]
        xs.append(features(coarse, bc))
        ys.append(np.stack([(a - b) * wet / s for a, b, s in zip(nxt, pred, osc)]).astype(np.float32))
    fine.close()
    coarse.close()
    return np.stack(xs), np.stack(ys)


def crops(x, y, size, count, rng):
    """Random size x size crops of samples (all samples at least that big)."""
    S, _, H, W = x.shape
    cx = np.empty((count, x.shape[1], size, size), np.float32)
    cy = np.empty((count, y.shape[1], size, size), np.float32)
    for k in range(count):
        s = rng.integers(S)
        i0, j0 = rng.integers(H - size + 1), rng.integers(W - size + 1)
        cx[k] = x[s, :, i0:i0 + size, j0:j0 + size]
        cy[k] = y[s, :, i0:i0 + size, j0:j0 + size]
    return cx, cy


def train(datasets, steps=2000, batch=16, size=64, lr=2e-3, seed=0, params=None, log_every=0):
    """Adam on random crops of the (x, y) datasets; returns (params, losses)."""
    rng = np.random.default_rng(seed)
    net = fwcnn.Net(size, size, batch, fwcnn.init_params(seed) if params is None else params)
    losses = []
    for it in range(steps):
        x, y = datasets[rng.integers(len(datasets))]
        cx, cy = crops(x, y, size, batch, rng)
        a = lr * 0.5 * (1 + np.cos(np.pi * it / steps))  # cosine decay
        loss, _ = net.train_step(cx, cy, a)
        losses.append(loss)
        if log_every and (it + 1) % log_every == 0:
            print(f"  train step {it + 1}: loss {np.mean(losses[-log_every:]):.4e}", flush=True)
    p = net.params()
    net.close()
    return p, losses


def run_fine(case, steps, precision="sp", state=None):
    """Plain run: final interior (eta, MASK) and device-complete wall
    seconds (from context creation; `state`, if given, is the case's
    prebuilt host state)."""
    st = state if state is not None else fwsetup.build_state(case, precision)
    t0 = time.perf_counter()
    g = fwgpu.Solver(st, precision)
    g.step(steps)
    g.synchronize()
    wall = time.perf_counter() - t0
    eta = interior(g.download("Eta")).astype(np.float64)
    mask = interior(g.download("MASK")).astype(np.float64)
    g.close()
    return (eta, mask), wall


def restricted_start(fine_state):
    """Interior Eta, U, V, Ubar, Vbar of a fine host state, restricted by
    restricted_fields: the coarse runs start from the fine initial condition
    (the reference's INI_SOL crest sits at a grid index, so each spacing
    would otherwise place it elsewhere)."""
    z = np.zeros_like(fine_state["Depth"], dtype=np.float64)
    return restricted_fields(lambda k: interior(np.asarray(fine_state.get(k, z), np.float64)))


def water_surface_error(fine_eta, fine_mask, coarse_eta, coarse_mask):
    """Relative L2 of the coarse water surface against the wet-cell mean of
    the fine one, over cells wet in the coarse run or in any cell of the
    fine block (0 elsewhere)."""
    ef, wf = restrict_wet(fine_eta, fine_mask)
    ec = coarse_eta * coarse_mask
    return float(np.linalg.norm(ec - ef) / max(np.linalg.norm(ef), 1e-300))


def water_volume(eta, wet, base):
    """Water volume of interior cells, sum over wet cells of (eta + depth)
    dx dy (conserved by the scheme away from sponges and sources)."""
    d = interior(base["Depth"]).astype(np.float64)
    return float(np.sum(np.where(wet, eta + d, 0.0))) * float(base["DX"]) * float(base["DY"])


def run_coarse(case, steps, params=None, precision="sp", start=None, state=None, conserve=False):
    """Coarse run of `steps` steps from `start` (restricted_start of the
    fine member) or the case's own initial state, corrected after every step
    when params is given (with conserve, a correction that moves no water
    volume: fwcnn.Net.set_conserve): final interior (eta, MASK), wall seconds
    the water volume at the end (water_volume) and the ratio of the
    energy at the end to that at the start (fw_reduce_stats)."""
    st = state if state is not None else fwsetup.build_state(case, precision)
    t0 = time.perf_counter()
    g = fwgpu.Solver(st, precision)
    if start is not None:
        set_state(g, st, *start)
    e_start = g.stats()["energy"]
    net = None
    if params is not None:
        net = fwcnn.Net(case.nglob, case.mglob, 1, params)
        net.set_conserve(conserve)
    for _ in range(steps):
        g.step(1)
        if net is not None:
            net.correct(g)
    g.synchronize()
    wall = time.perf_counter() - t0
    eta = interior(g.download("Eta")).astype(np.float64)
    mask = interior(g.download("MASK")).astype(np.float64)
    e_end = g.stats()["energy"]
    if net is not None:
        net.close()
    g.close()
    return (eta, mask), wall, water_volume(eta, mask > 0, st), e_end / e_start if e_start > 0 else float("inf")


FALLBACK_DRIFT = 1e-4
FALLBACK_ENERGY = 1.1


def compare(members, fine_steps, params, precision="sp", workers=1, fallback_drift=FALLBACK_DRIFT,
            fallback_energy=FALLBACK_ENERGY, conserve=True):
    """Per member: plain fine, plain coarse and learned coarse runs over the
    same simulated time; errors against the restricted fine surface and the
    runs/hour of each mode.  Host states are built first; then each mode runs
    all members, `workers` at a time from host threads (each member its own
    context and streams), and its runs/hour is members / that batch's wall
    time.  A learned member whose water volume drifts by more than
    `fallback_drift` of itself, or whose energy grows by more than the
    factor `fallback_energy` (the scheme only dissipates it), falls back to
    the fine run: its result is the fine one, and the learned mode's time
    includes a batch of those fine runs.
    With conserve the learned runs apply the correction without its
    wet-cell mean of eta (run_coarse)."""
    fine_st = [fwsetup.build_state(c, precision) for c in members]
    coarse = [coarse_case(c) for c in members]
    coarse_st = [fwsetup.build_state(c, precision) for c in coarse]
    starts = [restricted_start(s) for s in fine_st]
    idx = range(len(members))
    fwgpu.load_library()          # bind once, before any worker thread does
    fwcnn._lib()

    def batch(fn, which=idx):
        t0 = time.perf_counter()
        if workers <= 1:
            res = [fn(k) for k in which]
        else:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(workers) as ex:
                res = list(ex.map(fn, which))
        return res, time.perf_counter() - t0

    fine, t_f = batch(lambda k: run_fine(members[k], fine_steps, precision, fine_st[k]))
    plain, t_c = batch(lambda k: run_coarse(coarse[k], fine_steps // 2, None, precision, starts[k], coarse_st[k]))
    learned, t_l = batch(lambda k: run_coarse(coarse[k], fine_steps // 2, params, precision, starts[k],
                                              coarse_st[k], conserve))
    # water volume of each coarse start; drift = |V(end) - V(start)| / V(start)
    vol = [water_volume(starts[k][0], starts[k][0] > -interior(coarse_st[k]["Depth"]), coarse_st[k]) for k in idx]
    drift = lambda res, k: abs(res[k][2] - vol[k]) / vol[k]  # noqa: E731
    fb = [k for k in idx if not (drift(learned, k) <= fallback_drift and learned[k][3] <= fallback_energy)]
    t_fb = batch(lambda k: run_fine(members[k], fine_steps, precision, fine_st[k]), fb)[1] if fb else 0.0
    rows = []
    for k, case in enumerate(members):
        (ef, mf), wf = fine[k]
        (ec, mc), wc, _, gc = plain[k]
        (el, ml), wl, _, gl = learned[k]
        back = k in fb
        err_l = water_surface_error(ef, mf, el, ml)
        rows.append({"member": case.name, "M": case.mglob, "N": case.nglob,
                     "err_coarse": water_surface_error(ef, mf, ec, mc),
                     # the corrected coarse run's own error, fallback or not
                     "err_learned": err_l,
                     # the error of the result the learned mode delivers: the
                     # fine run's own (0) for a member that fell back
                     "err_delivered": 0.0 if back else err_l,
                     "drift_coarse": drift(plain, k), "drift_learned": drift(learned, k),
                     "energy_ratio_coarse": gc, "energy_ratio_learned": gl, "fallback": back,
                     "wall_fine": wf, "wall_coarse": wc, "wall_learned": wl})
    n = len(members)
    kept = [r for r in rows if not r["fallback"]]
    return {"members": rows, "workers": workers, "fallbacks": len(fb),
            "runs_per_hour": {"fine": 3600.0 * n / t_f, "coarse": 3600.0 * n / t_c,
                              "learned": 3600.0 * n / (t_l + t_fb)},
            # every delivered learned result within the bound, and at least
            # one member actually carried by the correction
            "all_learned_within_bound": bool(kept) and all(r["err_delivered"] <= ERR_BOUND for r in rows),
            "max_err_learned": max(r["err_learned"] for r in rows),
            "max_err_learned_kept": max((r["err_learned"] for r in kept), default=float("nan")),
            "members_improved": sum(r["err_learned"] < r["err_coarse"] for r in rows),
            "max_err_coarse": max(r["err_coarse"] for r in rows)}


def _batch_run(states, steps, precision, nets=None, conserve=False):
    """Advance the members of `states` together for `steps` steps through
    one fw_batch (one launch per phase over all members); with `nets` (one
    fwcnn.Net per member) the learned correction follows every step.
    Returns per member (eta, mask, water volume, energy ratio), and the wall
    seconds from context creation to the last step's completion."""
    t0 = time.perf_counter()
    gs = [fwgpu.Solver(st, precision) for st in states]
    return _batch_advance(gs, states, steps, nets, conserve, t0)


def _batch_advance(gs, states, steps, nets, conserve, t0):
    e0 = [g.stats()["energy"] for g in gs]
    b = fwgpu.Batch(gs)
    if nets is None:
        b.step(steps)
    else:
        for n in nets:
            n.set_conserve(conserve)
        for _ in range(steps):
            b.step(1)
            for g, n in zip(gs, nets):
                n.correct(g)
    for g in gs:
        g.synchronize()
    wall = time.perf_counter() - t0
    out = []
    for g, st, e in zip(gs, states, e0):
        eta = interior(g.download("Eta")).astype(np.float64)
        mask = interior(g.download("MASK")).astype(np.float64)
        e1 = g.stats()["energy"]
        out.append((eta, mask, water_volume(eta, mask > 0, st), e1 / e if e > 0 else float("inf")))
    b.close()
    for g in gs:
        g.close()
    return out, wall


def compare_batched(members, fine_steps, params, precision="sp", fallback_drift=FALLBACK_DRIFT,
                    fallback_energy=FALLBACK_ENERGY, conserve=True):
    """Per member: plain fine, plain coarse and learned coarse runs over the
    same simulated time, each mode's members advanced together as one
    ensemble (fw_batch); errors against the restricted fine surface and the
    runs/hour of each mode (members / the mode's wall time, from context
    creation to the last step).  A learned member whose water volume drifts
    by more than `fallback_drift` or whose energy grows by more than the
    factor `fallback_energy` falls back to the fine run: its delivered
    result is the fine one, and the learned mode's time includes a batch of
    those members' fine runs."""
    fine_st = [fwsetup.build_state(c, precision) for c in members]
    coarse = [coarse_case(c) for c in members]
    coarse_st = [fwsetup.build_state(c, precision) for c in coarse]
    starts = [restricted_start(s) for s in fine_st]
    fine, t_f = _batch_run(fine_st, fine_steps, precision)

    def coarse_solvers():
        gs = []
        for st, start in zip(coarse_st, starts):
            g = fwgpu.Solver(st, precision)
            set_state(g, st, *start)
            gs.append(g)
        return gs

    t0 = time.perf_counter()
    plain, t_c = _batch_advance(coarse_solvers(), coarse_st, fine_steps // 2, None, False, t0)
    t0 = time.perf_counter()
    nets = [fwcnn.Net(c.nglob, c.mglob, 1, params) for c in coarse]
    learned, t_l = _batch_advance(coarse_solvers(), coarse_st, fine_steps // 2, nets, conserve, t0)
    for n in nets:
        n.close()
    vol = [water_volume(starts[k][0], starts[k][0] > -interior(coarse_st[k]["Depth"]), coarse_st[k])
           for k in range(len(members))]
    drift = lambda res, k: abs(res[k][2] - vol[k]) / vol[k]  # noqa: E731
    fb = [k for k in range(len(members))
          if not (drift(learned, k) <= fallback_drift and learned[k][3] <= fallback_energy)]
    t_fb = _batch_run([fine_st[k] for k in fb], fine_steps, precision)[1] if fb else 0.0
    rows = []
    for k, case in enumerate(members):
        ef, mf = fine[k][0], fine[k][1]
        err_c = water_surface_error(ef, mf, plain[k][0], plain[k][1])
        err_l = water_surface_error(ef, mf, learned[k][0], learned[k][1])
        back = k in fb
        rows.append({"member": case.name, "M": case.mglob, "N": case.nglob,
                     "err_coarse": err_c, "err_learned": err_l, "err_delivered": 0.0 if back else err_l,
                     "drift_coarse": drift(plain, k), "drift_learned": drift(learned, k),
                     "energy_ratio_coarse": plain[k][3], "energy_ratio_learned": learned[k][3],
                     "fallback": back})
    n = len(members)
    kept = [r for r in rows if not r["fallback"]]
    return {"members": rows, "fallbacks": len(fb),
            "wall_s": {"fine": t_f, "coarse": t_c, "learned": t_l, "fallback_fine": t_fb},
            "runs_per_hour": {"fine": 3600.0 * n / t_f, "coarse": 3600.0 * n / t_c,
                              "learned": 3600.0 * n / (t_l + t_fb)},
            "all_learned_within_bound": bool(kept) and all(r["err_delivered"] <= ERR_BOUND for r in rows),
            "max_err_learned": max(r["err_learned"] for r in rows),
            "max_err_learned_kept": max((r["err_learned"] for r in kept), default=float("nan")),
            "members_improved": sum(r["err_learned"] < r["err_coarse"] for r in rows),
            "max_err_coarse": max(r["err_coarse"] for r in rows)}
