"""Benchmark: batched SCFT solves/s to ||phi - phi_t||_inf <= 1e-6 on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config5|config3]
                    [--n-inst N]

Workloads (independent instances, Dirichlet wall at x = 0, zero-flux end at
L, EXP wall target, w0 = 0.1 exp(-x/0.2) + N(0, 1e-3) per instance seed):

  config5 (default)  the production mix: 16384 instances on meshes graded
                     toward the wall, h(x) = h0 (1 + x/ell), h0 log-uniform
                     in [1e-4, 1e-2], ell uniform in [0.05, 0.5], 64-4096
                     nodes, N uniform over {512, ..., 8192}, L in [2, 8]
  config3            4096 uniform meshes, m log-uniform in [32, 4096], N in
                     {512, 1024, 2048, 4096}, L in [2, 8]

A step is one complete batched solve of every instance with the fields
already resident in HBM: the device Anderson mixer drives each instance
until its residual is below 1e-6.  Two legs are timed, K steps each:

  cold   every instance starts from w0 (the reference's initial guess)
  warm   every instance starts from the prediction of the warm-start MLP,
         trained (untimed setup) on fields that the batched solver itself
         converged for instances of the same workload drawn from disjoint
         seeds; the timed step includes the prediction (network forward and
         the mapping of its outputs onto every mesh, on the device, straight
         into the resident field); an instance whose first residual exceeds
         10x the batch median restarts from w0

`value` = warm-start instances solved per second (cold with --no-warm);
config.cold_start holds the cold leg and config.warm_over_cold the ratio.

Also reported:
  audit         64 random instances of the warm solve: the device's density
                of its converged field against the CPU oracle's evaluation of
                the same field, and the oracle's residual of it
  roofline      the dominant march launch (largest per-bucket time) of one
                evaluation sweep: algorithmic bytes (q-slabs written and read
                back, O(n) I/O) over its HIP-event duration, against 8 TB/s
  cpu_baseline  the CPU oracle (same mixer, same tolerance, cold start) on a
                bounded sample of the same instances, threads over the host
                cores the GPU box allots to one GPU; beside it the reference
                problem of configuration 1 (m = 32, N = 2048, w0 = 0, plain
                Anderson, history 10, mixing 0.1) solved single-threaded
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCFT solves/s to resid 1e-6, warm-start vs cold-start; propagator HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# The mixer of both sides: history 16 (the fine meshes' wall layer has many
# slowly converging modes: history 10 needs up to 32 evaluations on the
# uniform-mesh batch, 16 needs 25), log(phi/phi_t) mixing residual
# preconditioned by I + a M_L^{-1} K, beta 1, no step cap, a NaN evaluation
# halves the step (backtrack factor 1e30: NaN only), and the field bounded
# below by -0.5/dt (1 + dt w >= 1/2: the implicit-Euler matrix stays
# diagonally dominant; without it ~10% of the graded meshes never converge).
# a = 0.5 on the uniform meshes (their finest 48: 20.1 mean / 21 max
# evaluations against 20.8 / 22 at 0.6); a = 0.6 on the graded ones (their
# slowest 8: 30 / 36 against 83 / 101 at 0.5; 40 stiff ones 30 / 41 against
# 34 / 58).
WORKLOADS = {
    "config5": dict(
        n_inst=16384, seed=20261205, precond=0.6, node_weight=10.0, cuts=(512,),
        range_epochs=(1, 1), warm_inner=1,
        text="config5: {n} DN instances on graded meshes h(x) = h0 (1 + x/ell), h0 "
             "log-uniform [1e-4,1e-2], ell in [0.05,0.5], 64-4096 nodes, N in "
             "{{512,...,8192}}, L in [2,8], tol 1e-6"),
    "config3": dict(
        n_inst=4096, seed=20261202, precond=0.5, node_weight=10.0, cuts=(256,),
        range_epochs=(2, 1), warm_inner=4,
        text="config3: {n} DN instances, m log-uniform [32,4096], N in {{512,1024,2048,4096}}, "
             "L in [2,8], tol 1e-6"),
}


# The ground-state operator G = -(1/d^2)(d^2 f')' + c w+ (d = sqrt(phi_t),
# c = 1/4) preconditions every leg, in its rational form
#   (1 - a) f + a G f + a g (g I + G0)^-1 f,  a = 1/2, g = 3
# (G0: G without the field term), whose symbol follows the inverse Debye
# function 1/g_D(x) = x^2 / (2 (e^-x + x - 1)) to ~2.5%: the preconditioned
# Jacobian's eigenvalues lie within ~0.03 of 1 near the solution.  From w0
# the first updates are far from it: there the linear form f + a G f is
# used until an instance's ||r||_inf first falls below 0.5, then the
# rational one with the history cleared.
# Whole production-mix batches, seconds per solve: from w0, linear form
# throughout (a = 0.45) 1.058, rational after ||r|| < 0.5 / 0.3 / 0.1:
# 0.769 / 0.785 / 0.805 (a = 0.45: 0.872); from predictions, rational from
# the first update 0.286 against 0.339 linear (g = 2 or 4: 0.30-0.31).
# Field weight c from w0 (linear form): 0, 0.25, 0.5, 1 -> 1.17, 1.06, 1.25,
# 2.32 s (1 leaves a tail to 110 evaluations).
GS_PRECOND = 0.5
GS_SHIFT = 3.0
GS_FIELD_WEIGHT = 0.25
COLD_SHIFT_BELOW = 0.5
# Predicted starts: an instance not converged after 24 evaluations goes over
# to the plain form with its current field (history cleared), a safety net
# that no instance of either bench mix reaches.
GS_SWITCH = 24
# Predicted starts need a shorter history than w0: 8 columns.
WARM_HISTORY = 8
# The network: 40 000 converged fields, Adam in four stages of 100 epochs
# with the rate stepped down 1e-3 -> 3e-5.  A single 50-epoch stage at 1e-3
# leaves networks whose warm solves vary with the shuffling seed (production
# mix: mean 8.8-11.5 evaluations, 1.09-1.38 s per solve over five seeds and
# two data sets); the staged schedule gives 8.3-8.4 and ~1.0-1.2 s.
TRAIN_SCHEDULE = ((100, 1e-3), (100, 3e-4), (100, 1e-4), (100, 3e-5))
# The loss weights the wall node samples by node_weight^2 (WORKLOADS): on
# the graded meshes of the production mix the first residual of a
# prediction is set mostly there.  Warm solve there: weight 1, 3, 10, 30,
# 100 -> 0.270, 0.249, 0.234, 0.242, 0.278 s (median first residual 7.9e-5
# -> 5.1e-5 at 10).  On the uniform meshes of configuration 3, whose coarse
# instances have no node samples, weight 1 was best for one network; with
# the networks split by mesh size (below) weight 1 / 3 / 10 -> 35.1-35.7 /
# 33.6-34.1 / 33.9-34.1 ms (10 kept: the lowest first residuals, 2.92 mean
# evaluations); on the production mix split, 3 / 10 / 30 -> 234 / 228 / 244.
# One network per mesh-size range (cuts in WORKLOADS, element counts): warm
# solve of the production mix one network / split at 256 / 512 / 1024 ->
# 238 / 233 / 228 / 233 ms; of configuration 3 one / split at 128 / 256 /
# 512 -> 39-45 / 36-37 / 35-36 / 36 ms (median first residual of the
# m >= 512 instances 1.3e-4 -> 3.7e-5 at 256).  range_epochs multiplies a
# range's epochs: configuration 3's coarse network (m < 256) at twice the
# schedule -> 2.62 instead of 2.92 mean evaluations, 139 instead of 386
# fallbacks, 30.5 ms per warm step; the fine networks gain nothing from more.
# Inner sweeps (warm_inner in WORKLOADS: evaluations per sweep of the
# instances outside the bucket of most work; the same results, another
# schedule): configuration 3's warm solve with 1 / 2 / 4 / 8 -> 35.0-38.8 /
# 34.1-35.0 / 32.8-33.8 / 36.8-37.4 ms (9 -> 5 sweeps: the coarse meshes'
# extra evaluations and fallbacks run beside the large instances' march);
# the production mix's warm solve and both cold legs gain nothing (the
# largest bucket's own evaluations set their sweep count) or lose to the
# extra concurrent traffic (configuration 3 cold: 82 -> 87 ms at 4).


def solver_settings(workload):
    """The solver from w0 (the cold leg, the training solves, the CPU
    baseline's cold pair); predicted starts change it in warm_overrides."""
    return dict(history=16, beta=1.0, tol=1e-6, step_cap=0.0, floor=1e-8,
                precond=WORKLOADS[workload]["precond"], backtrack=1e30, max_halvings=30,
                field_floor=0.5, gs_precond=GS_PRECOND, gs_field_weight=GS_FIELD_WEIGHT,
                gs_shift=GS_SHIFT, gs_shift_below=COLD_SHIFT_BELOW)


def warm_overrides():
    """From a predicted start: a shorter history and the plain form after
    GS_SWITCH evaluations.  The threshold rule stays: a prediction's first
    residual is far below it, so the rational form applies from the first
    update, while an instance that falls back to w0 starts linear as the
    cold leg does."""
    return dict(history=WARM_HISTORY, gs_switch=GS_SWITCH)


def solver_text(workload):
    return (f"device Anderson, history 16 (cold) or {WARM_HISTORY} (predicted starts), log "
            f"mixing preconditioned by (1 - {GS_PRECOND}) I + {GS_PRECOND} G + "
            f"{GS_PRECOND}*{GS_SHIFT} ({GS_SHIFT} I + G0)^-1, G = -(1/d^2)(d^2 f')' + "
            f"{GS_FIELD_WEIGHT} w+, d = sqrt(phi_t), G0 = G without w+ (from w0: I + "
            f"{GS_PRECOND} G until ||r|| < {COLD_SHIFT_BELOW}; predicted starts: plain form I + "
            f"{WORKLOADS[workload]['precond']} M^-1 K after {GS_SWITCH} evaluations), beta 1.0, "
            "no step cap, NaN backtracking, field bounded below by -0.5/dt")


def make_workload(WL, workload, n_inst, seed=None):
    seed = WORKLOADS[workload]["seed"] if seed is None else seed
    if workload == "config5":
        return WL.config5(n_inst=n_inst, seed=seed)
    return WL.config3(n_inst=n_inst, seed=seed)


def march_source_sha256():
    """Hash of the march kernel's sources (the kernel and the device helpers
    it inlines: slab layout and access, lane moves)."""
    import hashlib
    h = hashlib.sha256()
    for name in ("march.hip", "device_common.h"):
        with open(os.path.join(ROOT, "scft_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def measured_traffic(workload, kern, n_items, alg_bytes):
    """HBM bytes of the dominant launch from the committed counter runs
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), when they were taken on this
    same launch (kernel source, variant, bucket, algorithmic bytes);
    otherwise None."""
    path = os.path.join(ROOT, "profiles", f"march_{workload}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("source_sha256") != march_source_sha256() or t.get("kernel") != kern or \
            t.get("items") != n_items or \
            abs(t.get("algorithmic_bytes", -1.0) - alg_bytes) > 1e-6 * alg_bytes:
        return None
    return round(float(t["traffic_bytes"]))


def warm_start_leg(S, WL, b, cfg, w0, args):
    """Train the warm-start network on fields the batched solver converges
    for `args.train_inst` instances drawn like the bench batch from disjoint
    seeds (setup, untimed), then time K warm-started solves of the bench
    batch.  A timed step covers the prediction (network forward and the
    mapping of its outputs onto every mesh, both on the device, writing the
    batch's resident field) and the device-resident solve;
    an instance whose first residual exceeds 10x the batch median restarts
    from w0."""
    from scft_amd import warmstart as WS
    if args.net_file and os.path.exists(args.net_file):
        ws = WS.WarmStart(S, seed=11, cuts=WORKLOADS[args.workload]["cuts"])
        rec = ws.load(args.net_file)
        log(f"bench: warm start: network loaded from {args.net_file} ({rec})")
        return warm_solves(S, b, cfg, w0, args, ws, rec)
    t0 = time.perf_counter()
    Xs, Ys = [], []
    opts = solver_opts(S, args)
    left, part = args.train_inst, 0
    while left > 0:
        n = min(8192, left)
        # seeds disjoint from the bench batch's (parameters and noise)
        tcfg = make_workload(WL, args.workload, n, seed=7000 + 101 * part)
        out = WS.solve(S, tcfg, opts)
        conv = (out["status"] & S.ST_CONVERGED) != 0
        X, Y = WS.dataset(S, tcfg, out["w_nodes"], conv)
        Xs.append(X)
        Ys.append(Y)
        left -= n
        part += 1
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    t_data = time.perf_counter() - t0
    ws = WS.WarmStart(S, seed=11, cuts=WORKLOADS[args.workload]["cuts"],
                      range_epochs=WORKLOADS[args.workload]["range_epochs"])
    t0 = time.perf_counter()
    hist = ws.fit(X, Y, batch=256, seed=12, schedule=TRAIN_SCHEDULE,
                  node_weight=WORKLOADS[args.workload]["node_weight"])
    t_train = time.perf_counter() - t0
    hists = hist if isinstance(hist, list) else [hist]
    log(f"bench: warm start: {len(X)} training fields in {t_data:.1f} s, trained in "
        f"{t_train:.1f} s: " + "; ".join(f"{len(h)} epochs, loss {h[0]:.3e} -> {h[-1]:.3e}"
                                          for h in hists))
    rec = {"training_fields": int(len(X)), "epochs": int(max(len(h) for h in hists)),
           "final_loss": round(float(np.mean([h[-1] for h in hists])), 6),
           "train_seconds": round(t_train, 2)}
    if args.net_file:
        ws.save(args.net_file, **rec)
    return warm_solves(S, b, cfg, w0, args, ws, rec)


def warm_solves(S, b, cfg, w0, args, ws, rec):
    """Time K warm-started solves of the bench batch with the trained
    network `ws` (its training record `rec`)."""
    from scft_amd import warmstart as WS
    prep = WS.Prepared(S, cfg)
    wopts = solver_opts(S, args)
    wopts.fallback_factor = 10.0
    for k, v in warm_overrides().items():  # from a predicted start
        setattr(wopts, k, v)
    wopts.inner_sweeps = WORKLOADS[args.workload]["warm_inner"]
    w0_flat = b.join(w0)
    b.upload(w0_flat, w0_flat)  # the cold fields the fallback restarts from, resident in HBM
    plan = ws.device_plan(b, prep)  # network inputs and mapping tables, resident in HBM

    def step():
        t = time.perf_counter()
        plan.predict()  # forward + mapping kernel: writes the batch's field in HBM
        t_pred = time.perf_counter() - t
        r = b.solve_resident(wopts)
        return t_pred, r

    for i in range(args.warmup):
        t_pred, r = step()
        log(f"bench: warm warmup {i}: {t_pred + r['seconds']:.3f} s")
    total, res, tp = 0.0, None, 0.0
    for i in range(args.steps):
        t_pred, r = step()
        total += t_pred + r["seconds"]
        tp += t_pred
        res = b.download()
        fb = int(np.sum((res["status"] & S.ST_FALLBACK) != 0))
        log(f"bench: warm step {i}: {t_pred:.3f} s predict + {r['seconds']:.3f} s solve, "
            f"{r['sweeps']} sweeps, {res['converged']}/{len(cfg['xs'])} converged, mean iters "
            f"{float(np.mean(res['iters'])):.1f}, max {int(np.max(res['iters']))}, {fb} fallbacks")
    # first residuals of the predicted and the cold fields (outside the timed steps)
    plan.predict()
    w_pred = b.download()["w"]
    first_warm = b.eval(w_pred)["rnorm"]
    first_cold = b.eval(w0)["rnorm"]
    plan.close()
    q = lambda a: [float(f"{v:.3g}") for v in np.quantile(a, [0.1, 0.5, 0.9])]
    log(f"bench: first residual quantiles 10/50/90%: warm {q(first_warm)}, cold {q(first_cold)}")
    ws.close()
    info = {
        "first_residual_q10_q50_q90": q(first_warm),
        "cold_first_residual_q10_q50_q90": q(first_cold),
        "network": f"{WS.DIMS[0]} -> 4 x 256 tanh -> {WS.DIMS[-1]}, FP32, device "
                   "forward/backward/Adam kernels; one per mesh-size range, split at m = "
                   + ", ".join(str(c) for c in WORKLOADS[args.workload]["cuts"]),
        "training_fields": rec.get("training_fields"), "epochs": rec.get("epochs"),
        "batch": 256, "lr_schedule": [list(st) for st in TRAIN_SCHEDULE],
        "node_weight": WORKLOADS[args.workload]["node_weight"],
        "inner_sweeps": WORKLOADS[args.workload]["warm_inner"],
        "range_epochs": list(WORKLOADS[args.workload]["range_epochs"]),
        "final_loss": rec.get("final_loss"), "train_seconds": rec.get("train_seconds"),
        "predict_ms_per_step": round(1e3 * tp / args.steps, 2),
        "fallbacks": int(np.sum((res["status"] & S.ST_FALLBACK) != 0)),
    }
    return dict(res=res, total=total, value=res["converged"] * args.steps / total, info=info,
                starts=[np.concatenate([[0.0], v]) for v in w_pred])


def solver_opts(S, args):
    return S.anderson_opts(max_iter=args.max_iter, transform=S.MIX_LOG, check_every=4,
                           progress_seconds=20.0, **solver_settings(args.workload))


def continuation_compare(args):
    """The reference's own warm start, mesh continuation (DEALII_SCFT/
    drivescft.cc:291-322: solve on a coarser mesh, carry the field to the
    finer one by spline, solve there), on the bench batch: every mesh is
    coarsened to every other node (the mesh whose uniform bisection gives it
    back), the coarse batch is solved from w0, its fields are moved onto the
    bench meshes by the not-a-knot spline (scft_resample_spline_batch, as
    spline_chen.c), and the bench batch is solved from them with the cold
    fallback.  Timed per step: both solves and the transfer.  The fine solve
    uses the MLP leg's settings for predicted starts (the transferred field
    is one).  The cold leg is timed beside it; the MLP leg is the bench
    line's own on the same batch."""
    import scft_amd as S
    from scft_amd import workloads as WL
    if args.n_inst <= 0:
        args.n_inst = WORKLOADS[args.workload]["n_inst"]
    cfg = make_workload(WL, args.workload, args.n_inst)
    xs = cfg["xs"]
    xc = [np.concatenate([x[:-1:2], x[-1:]]) if (len(x) - 1) % 2 else x[::2] for x in xs]
    xc = [x if len(x) >= 5 else xs[i] for i, x in enumerate(xc)]
    w0 = [w[1:] for w in cfg["w0"]]
    w0c = [WL.w0_field(x, WORKLOADS[args.workload]["seed"] + 1 + k)[1:] for k, x in enumerate(xc)]
    fine = S.Batch(xs, cfg["Ns"], [p[1:] for p in cfg["pts"]], bc=S.BC_DN)
    coarse = S.Batch(xc, cfg["Ns"], [WL.exp_target(x)[1:] for x in xc], bc=S.BC_DN)
    for bb, w in ((fine, w0), (coarse, w0c)):
        bb.upload(w)
        bb.autotune(reps=2)
    opts = solver_opts(S, args)
    wopts = solver_opts(S, args)
    wopts.fallback_factor = 10.0
    for k, v in warm_overrides().items():  # a transferred field is a predicted start
        setattr(wopts, k, v)
    w0_flat = fine.join(w0)
    # the spline transfer: coarse unknowns (with the wall node) -> fine unknowns
    src_x = [x for x in xc]
    dst_x = [x[1:] for x in xs]

    def cont_step():
        t0 = time.perf_counter()
        coarse.upload(w0c)
        rc = coarse.solve_resident(opts)
        resc = coarse.download()
        t1 = time.perf_counter()
        src_y = [np.concatenate([[0.0], v]) for v in resc["w"]]
        wt = S.resample_spline_batch(src_x, src_y, dst_x, flat=True)[0]
        fine.upload(wt, w0_flat)
        t2 = time.perf_counter()
        rf = fine.solve_resident(wopts)
        t3 = time.perf_counter()
        return t3 - t0, (t1 - t0, t2 - t1, rf["seconds"]), resc, rc

    def cold_step():
        fine.upload(w0)
        r = fine.solve_resident(opts)
        return r["seconds"]

    for _ in range(args.warmup):
        cont_step()
        cold_step()
    tc, tw, parts = 0.0, 0.0, np.zeros(3)
    for i in range(args.steps):
        tw += cold_step()
        res_cold = fine.download()
        dt, pp, resc, rc = cont_step()
        res = fine.download()
        tc += dt
        parts += pp
        log(f"bench: continuation step {i}: coarse solve {pp[0]:.3f} s ({rc['sweeps']} sweeps, "
            f"mean iters {resc['iters'].mean():.1f}), transfer {pp[1]:.3f} s, fine solve "
            f"{pp[2]:.3f} s ({res['sweeps']} sweeps, mean iters {res['iters'].mean():.1f}, "
            f"max {res['iters'].max()}); cold {res_cold['iters'].mean():.1f} mean iters")
    out = {"workload": WORKLOADS[args.workload]["text"].format(n=args.n_inst),
           "cold_solves_per_s": round(args.n_inst * args.steps / tw, 1),
           "continuation_solves_per_s": round(args.n_inst * args.steps / tc, 1),
           "continuation_seconds_coarse_transfer_fine": [round(v / args.steps, 4) for v in parts],
           "continuation_fine_mean_iters": round(float(res["iters"].mean()), 2),
           "continuation_fine_max_iters": int(res["iters"].max()),
           "continuation_coarse_mean_iters": round(float(resc["iters"].mean()), 2),
           "cold_mean_iters": round(float(res_cold["iters"].mean()), 2),
           "converged": int(res["converged"])}
    fine.close()
    coarse.close()
    print(json.dumps(out), flush=True)


def stepper_compare(args):
    """Sweep time of 4096 copies of the field the reference's deal.II IRK
    solver converged (DEALII_SCFT/inputFiles/N=33_for_read.txt, consistent
    w-mass, wall values from the file), implicit Euler against IRK4, with the
    residual each leaves on that field (the discretisation error of phi)."""
    import re
    import scft_amd as S
    from scft_amd import workloads as WL
    path = os.path.join(ROOT, "tests", "golden", "files", "N=33_for_read.txt")
    rows = [l for l in open(path) if re.match(r"^\s*\d+\s*,", l)]
    d = np.array([[float(v) for v in r.split(",")] for r in rows])
    x, w = d[:, 1], d[:, 2]
    pt = WL.exp_target(np.minimum(x, x[-1] - x))  # both walls
    n = args.n_inst
    for stepper, N in [(S.STEP_IE, 4096), (S.STEP_IE, 2048), (S.STEP_IRK4, 256), (S.STEP_IRK4, 512)]:
        b = S.Batch([x] * n, [N] * n, [pt[1:-1]] * n, bc=S.BC_DD, cmat=S.C_CONSISTENT,
                    stepper=stepper)
        b.set_walls([w[0]] * n, [w[-1]] * n)
        b.upload([w[1:-1]] * n)
        b.autotune(reps=2)
        ms = b.time_eval([w[1:-1]] * n, reps=10)
        out = b.eval([w[1:-1]] * n)
        b.close()
        log(f"bench: {'IE' if stepper == S.STEP_IE else 'IRK4'} N={N}: {ms:.3f} ms per sweep of "
            f"{n} instances, ||r|| of the reference field {out['rnorm'][0]:.2e}")


def gap_curve(args):
    """First residual -> evaluations with the warm leg's solver, on the bench
    batch itself (device only): every instance converged from w0 to 1e-11,
    its field perturbed by a smooth relative bump of size eps,
    w (1 + eps (cos(pi x/L) + sin(3 pi x/L)/2)), and solved back to the
    bench tolerance from there (no fallback).  Shows how good a prediction
    has to be for a given number of evaluations.  One JSON line."""
    import scft_amd as S
    from scft_amd import workloads as WL
    if args.n_inst <= 0:
        args.n_inst = WORKLOADS[args.workload]["n_inst"]
    cfg = make_workload(WL, args.workload, args.n_inst)
    w0 = [w[1:] for w in cfg["w0"]]
    b = S.Batch(cfg["xs"], cfg["Ns"], [p[1:] for p in cfg["pts"]], bc=S.BC_DN)
    b.upload(w0)
    b.autotune(reps=2)
    tight = solver_opts(S, args)
    tight.tol = 1e-11
    b.upload(w0)
    b.solve_resident(tight)
    ref = b.download()
    log(f"bench: gap curve: {ref['converged']}/{args.n_inst} converged to 1e-11")
    wopts = solver_opts(S, args)
    for k, v in warm_overrides().items():
        setattr(wopts, k, v)
    bumps = [np.cos(np.pi * x[1:] / x[-1]) + 0.5 * np.sin(3.0 * np.pi * x[1:] / x[-1])
             for x in cfg["xs"]]
    rows = []
    for eps in (1e-2, 1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6):
        start = [w * (1.0 + eps * bb) for w, bb in zip(ref["w"], bumps)]
        first = b.eval(start)["rnorm"]
        ts = []
        for _ in range(2):
            b.upload(start)
            ts.append(b.solve_resident(wopts)["seconds"])
        res = b.download()
        it = np.asarray(res["iters"])
        row = {"eps": eps,
               "first_residual_q10_q50_q90": [float(f"{v:.3g}") for v in
                                               np.quantile(first, [0.1, 0.5, 0.9])],
               "mean_evaluations": round(float(it.mean()), 2), "max_evaluations": int(it.max()),
               "solve_seconds": round(min(ts), 4), "converged": int(res["converged"])}
        log(f"bench: gap curve {row}")
        rows.append(row)
    b.close()
    print(json.dumps({"workload": WORKLOADS[args.workload]["text"].format(n=args.n_inst),
                      "solver": solver_text(args.workload) + " (predicted-start settings)",
                      "curve": rows}), flush=True)


def cpu_baseline(cfg, gpu_iters, args, threads, budget_s, starts=None, overrides=None):
    """The CPU oracle (same mixer, options and tolerance) solving a random
    sample of the batch on `threads` OpenMP threads, from `starts` (fields on
    all nodes; default w0) with the solver settings plus `overrides`.  The
    sample is drawn until its estimated CPU work reaches `budget_s`; the
    full-batch CPU time is the sample's time scaled by the work ratio
    sum(iters * n * N)."""
    import oracle as O
    O.build()
    n_inst = len(cfg["xs"])
    work = np.array([gpu_iters[k] * (len(cfg["xs"][k]) - 1) * cfg["Ns"][k]
                     for k in range(n_inst)], dtype=np.float64)
    # seconds per node-step of one evaluation, from one mid-size evaluation
    k0 = int(np.argsort(work)[n_inst // 2])
    t0 = time.perf_counter()
    O.evaluate(cfg["xs"][k0], cfg["w0"][k0], cfg["pts"][k0], cfg["Ns"][k0], bc=O.BC_DN)
    c = (time.perf_counter() - t0) / ((len(cfg["xs"][k0]) - 1) * cfg["Ns"][k0])
    rng = np.random.default_rng(1)
    idx, est = [], 0.0
    for k in rng.permutation(n_inst):
        if est >= budget_s * threads and len(idx) >= threads:
            break
        idx.append(int(k))
        est += work[k] * c
    opts = O.anderson_opts(max_iter=args.max_iter, transform=O.MIX_LOG,
                           **{**solver_settings(args.workload), **(overrides or {})})
    starts = cfg["w0"] if starts is None else starts
    t0 = time.perf_counter()
    out = O.batch_solve([len(cfg["xs"][k]) for k in idx], [cfg["Ns"][k] for k in idx],
                        [cfg["xs"][k] for k in idx], [cfg["pts"][k] for k in idx],
                        [starts[k] for k in idx], opts, bc=O.BC_DN, threads=threads)
    dt = time.perf_counter() - t0
    sample_work = float(np.sum([out["iters"][i] * (len(cfg["xs"][k]) - 1) * cfg["Ns"][k]
                                for i, k in enumerate(idx)]))
    full_time = dt * float(work.sum()) / sample_work
    conv = int(np.sum(out["status"] == 0))
    return dict(value=n_inst / full_time, seconds=dt, converged=conv, sample=len(idx),
                full_time=full_time)


def config1_single_thread(S, WL):
    """The reference problem of configuration 1 (m = 32, N = 2048, w0 = 0,
    DN, L = 3.72) with the reference's plain Anderson mixing (history 10,
    mixing 0.1, step cap 0.5) to 1e-6: the CPU oracle on one thread, as the
    reference runs (1D_FEM.c:208-228), beside the device solving the same
    single instance."""
    import oracle as O
    cfg = WL.config1()
    x, N, pt, w0 = cfg["xs"][0], cfg["Ns"][0], cfg["pts"][0], cfg["w0"][0]
    opts = O.anderson_opts(history=10, beta=0.1, tol=1e-6, max_iter=5000, step_cap=0.5)
    reps, t0 = 0, time.perf_counter()
    while reps < 3 or time.perf_counter() - t0 < 2.0:
        cpu = O.solve_anderson(x, pt, N, w0, opts, bc=O.BC_DN)
        reps += 1
    t_cpu = (time.perf_counter() - t0) / reps
    b = S.Batch([x], [N], [pt[1:]], bc=S.BC_DN)
    dopts = S.anderson_opts(history=10, beta=0.1, tol=1e-6, max_iter=5000, step_cap=0.5,
                            check_every=8)
    b.solve([w0[1:]], dopts)  # warm-up (graph instantiation)
    t0 = time.perf_counter()
    dev = b.solve([w0[1:]], dopts)
    t_dev = time.perf_counter() - t0
    b.close()
    return {"workload": "config1: m=32, N=2048, L=3.72, w0=0, plain Anderson history 10, "
                        "mixing 0.1, tol 1e-6",
            "cpu_one_thread_seconds": round(t_cpu, 4),
            "cpu_evaluations": int(cpu["iterations"]),
            "cpu_converged": bool(cpu["rc"] == 0),
            "gpu_seconds": round(t_dev, 4),
            "gpu_evaluations": int(dev["iters"][0]),
            "gpu_converged": bool(dev["status"][0] & S.ST_CONVERGED)}


def audit(S, cfg, b, res, n_audit=64, threads=16):
    """Audit subset of a solve: the device's density of each converged field
    against the oracle's evaluation of the same field, and the oracle's
    residual of it."""
    import oracle as O
    out = b.eval(res["w"])
    rng = np.random.default_rng(64)
    idx = rng.choice(len(cfg["xs"]), n_audit, replace=False)
    wf = [np.concatenate([[0.0], res["w"][k]]) for k in idx]
    ref = O.batch_eval([len(cfg["xs"][k]) for k in idx], [cfg["Ns"][k] for k in idx],
                       [cfg["xs"][k] for k in idx], [cfg["pts"][k] for k in idx], wf,
                       bc=O.BC_DN, threads=threads)
    dphi = max(float(np.abs(out["phi"][k] - ref["phi"][j][1:]).max()) for j, k in enumerate(idx))
    return {"instances": int(n_audit), "max_abs_phi_dev_minus_oracle": dphi,
            "max_oracle_residual": float(np.max(ref["rnorm"])),
            "converged": int(np.sum((res["status"][idx] & S.ST_CONVERGED) != 0))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="config5")
    ap.add_argument("--n-inst", type=int, default=0,
                    help="instances (default: the workload's, 16384 / 4096)")
    ap.add_argument("--max-iter", type=int, default=500)
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU-baseline work (per thread) to time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-warm", action="store_true", help="cold start only")
    ap.add_argument("--net-file", default=None,
                    help="warm-start network file: loaded if it exists (no training), "
                         "else written after training")
    ap.add_argument("--no-cold", action="store_true",
                    help="skip the cold leg (profiling runs; no cold figures in the line)")
    ap.add_argument("--train-inst", type=int, default=40000,
                    help="converged fields the warm-start network is trained on")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the dominant-launch measurement (for profiler runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB 5 times and exit (counter scale)")
    ap.add_argument("--tune-file", default=None)
    ap.add_argument("--continuation-compare", action="store_true",
                    help="the reference's mesh continuation as the warm start of the bench "
                         "batch (coarse solve + spline transfer + solve), against the cold "
                         "start (one JSON line, not the bench line)")
    ap.add_argument("--gap-curve", action="store_true",
                    help="first residual -> evaluations of the warm leg's solver from "
                         "perturbed converged fields of the bench batch (one JSON line, not "
                         "the bench line)")
    ap.add_argument("--stepper-compare", action="store_true",
                    help="time one sweep of 4096 copies of the reference's 33-node deal.II "
                         "problem with implicit Euler and with IRK4 (no JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): instances are "
              "independent, so a multi-GPU run is N copies of this one", file=sys.stderr)
        sys.exit(2)

    if args.stepper_compare:
        stepper_compare(args)
        return
    if args.continuation_compare:
        continuation_compare(args)
        return
    if args.gap_curve:
        gap_curve(args)
        return

    if args.copy_calibration:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE per
        # dispatch calibrate the counters read for the march kernel
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copied {args.copy_calibration} MiB 5 times")
        return

    import scft_amd as S
    from scft_amd import workloads as WL

    if args.n_inst <= 0:
        args.n_inst = WORKLOADS[args.workload]["n_inst"]
    log(f"bench: generating {args.workload} ({args.n_inst} instances)")
    cfg = make_workload(WL, args.workload, args.n_inst)
    w0 = [w[1:] for w in cfg["w0"]]
    b = S.Batch(cfg["xs"], cfg["Ns"], [p[1:] for p in cfg["pts"]], bc=S.BC_DN,
                tune_file=args.tune_file)
    b.upload(w0)  # the field the tuning runs time
    t0 = time.perf_counter()
    if not args.no_autotune:  # variant choice per bucket: setup, outside the timed region
        b.autotune(reps=2)  # ends with the choice of the scratch allocation
        log(f"bench: autotune {time.perf_counter() - t0:.1f} s")
    else:  # variants from the tune file; the scratch allocation is chosen here
        b.tune_placement()
        log(f"bench: scratch placement chosen in {time.perf_counter() - t0:.1f} s")
    log(f"bench: buckets (n_max, K, W) = {b.tuning()}")

    # roofline of the dominant march launch (HIP events on the launch stream)
    buckets = b.time_buckets(w0, reps=3)
    tuning = b.tuning()  # same bucket order as time_buckets
    di = max(range(len(buckets)), key=lambda i: buckets[i][1])
    n_items, ms, by = buckets[di]
    kern = f"march_kernel<{tuning[di][1]}, {tuning[di][2]}>"
    achieved = by / (ms * 1e-3) / 1e9
    sweep_ms = sum(t[1] for t in buckets)
    log(f"bench: buckets (items, ms, GB) = {[(i, round(m, 3), round(bb / 1e9, 3)) for i, m, bb in buckets]}")
    log(f"bench: one sweep {sweep_ms:.2f} ms, dominant launch {ms:.3f} ms, {achieved:.1f} GB/s")
    if args.roofline_only:
        print(json.dumps({"roofline": {"achieved": round(achieved, 1), "ms": ms, "bytes": by,
                                       "items": n_items, "kernel": kern}}), flush=True)
        b.close()
        return

    # ---- cold start: every instance from w0 (the reference's initial guess)
    opts = solver_opts(S, args)
    if args.no_cold and args.no_warm:
        sys.exit("bench.py: --no-cold with --no-warm leaves nothing to time")
    for i in range(0 if args.no_cold else args.warmup):
        b.upload(w0)
        r = b.solve_resident(opts)
        log(f"bench: cold warmup {i}: {r['seconds']:.2f} s, {r['sweeps']} sweeps")
    total_cold = 0.0
    res_cold = None
    for i in range(0 if args.no_cold else args.steps):
        b.upload(w0)
        r = b.solve_resident(opts)
        total_cold += r["seconds"]
        res_cold = b.download()
        log(f"bench: cold step {i}: {r['seconds']:.3f} s, {r['sweeps']} sweeps, "
            f"{res_cold['converged']}/{args.n_inst} converged, mean iters "
            f"{float(np.mean(res_cold['iters'])):.1f}, max {int(np.max(res_cold['iters']))}")
    value_cold = res_cold["converged"] * args.steps / total_cold if res_cold else None

    # ---- warm start: the MLP trained on fields the batched solver produced
    warm = None
    if not args.no_warm:
        warm = warm_start_leg(S, WL, b, cfg, w0, args)
    res = warm["res"] if warm else res_cold
    total = warm["total"] if warm else total_cold
    value = warm["value"] if warm else value_cold
    gb_eff = res["bytes"] / (total / args.steps) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload]["text"].format(n=args.n_inst),
            "n_inst": args.n_inst,
            "start": "warm (MLP) with cold fallback" if warm else "cold",
            "solver": solver_text(args.workload),
            "converged": int(res["converged"]),
            "mean_iters": round(float(np.mean(res["iters"])), 2),
            "max_iters": int(np.max(res["iters"])),
            "solve_gbs_algorithmic": round(gb_eff, 1),
            "cold_start": None if res_cold is None else {
                "value": round(value_cold, 3), "unit": "solves/s",
                "ms_per_step": round(1e3 * total_cold / args.steps, 3),
                "converged": int(res_cold["converged"]),
                "mean_iters": round(float(np.mean(res_cold["iters"])), 2),
                "max_iters": int(np.max(res_cold["iters"])),
            },
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.workload, kern, n_items, by),
            "kernel": f"{kern}: bucket of {n_items} instances, {ms:.3f} ms per launch",
        },
    }
    if warm:
        out["config"]["warm_start"] = warm["info"]
        out["config"]["warm_over_cold"] = round(value / value_cold, 3) if value_cold else None
    out["config"]["audit"] = audit(S, cfg, b, res)
    log(f"bench: audit {out['config']['audit']}")
    if not args.no_cpu and res_cold is not None:
        # the host cores the GPU box allots to one GPU (OMP_NUM_THREADS there;
        # the whole host may show more): every one of them
        affinity = len(os.sched_getaffinity(0))
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
        threads = max(1, min(threads, affinity))
        def sample_text(c, what):
            return (f"{c['sample']} random instances of the same batch solved by the CPU "
                    f"oracle from {what} (same mixer, options, tolerance) in "
                    f"{c['seconds']:.1f} s on {threads} threads, {c['converged']} converged; "
                    f"full batch {c['full_time']:.0f} s by work ratio sum(iters*n*N)")
        cbc = cpu_baseline(cfg, res_cold["iters"], args, threads,
                           args.cpu_budget / 2 if warm else args.cpu_budget)
        if warm:  # the measured workload: the same predicted fields, the warm leg's options
            cb = cpu_baseline(cfg, res["iters"], args, threads, args.cpu_budget,
                              starts=warm["starts"],
                              overrides=warm_overrides())
            what = "the network's predicted fields (no fallback)"
        else:
            cb, what = cbc, "w0"
        c1 = config1_single_thread(S, WL)
        out["cpu_baseline"] = {
            "value": round(cb["value"], 4), "unit": "solves/s", "cores": threads,
            "kind": "port", "sample": sample_text(cb, what),
            "host_cpu_count": os.cpu_count(), "host_affinity": affinity,
            "gpu_over_cpu": round(value / cb["value"], 1),
            "cold": {"value": round(cbc["value"], 4), "unit": "solves/s",
                     "sample": sample_text(cbc, "w0"),
                     "gpu_over_cpu": round(value_cold / cbc["value"], 1)},
            "config1_single_thread": c1,
        }
        log(f"bench: cpu baseline {cb}; cold {cbc}; config 1 {c1}")
    b.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
