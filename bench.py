"""Benchmark: batched FWI gradients (forward + checkpointed adjoint) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--shots 64] [--nt 4000]

Workload: batched isotropic elastic FWI gradients -- 64 shots on a 500 x 1700 grid
(h = 10 m), nt = 4000 steps of 1 ms, 100-400 split-spread receivers per shot, 8-20 Hz
Ricker sources, free surface on top, 30-cell tapers elsewhere.  A "step" is one
gradient evaluation of the whole batch: forward modelling with device checkpoints, the
data residual, and the adjoint propagation with the imaging condition.  Inputs
(model, source traces, observed data) are resident in HBM before timing starts.

Printed JSON line: shots/s (value) and Gcell-updates/s of the batched form, the plain
one-shot-per-launch form on the same kernels, the roofline of the dominant kernel (the
fused time step) from HIP events, the gradient error of a reduced case against the
FP64 oracle, and a CPU baseline (the oracle's FP32 restatement on the host cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FWI gradient shots/s & Gcell-updates/s, batched vs plain, grad err≤1e-4"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE_FWD = 36.0   # iso forward step: read curr 2 + prev 2 + model 3, write next 2 floats
# one gradient, per cell and time step: the forward step (36 B iso, 44 B VTI) plus the adjoint
# step with imaging at sub = 1 (68 B iso, 76 B VTI: + 8 B of forward state, + 24 B read-modify-
# write of the 3 gradient fields); recomputation of the forward wavefield counts as overhead
BYTES_PER_GRADIENT_CELL_STEP = {False: 36.0 + 68.0, True: 44.0 + 76.0}
# acoustic operator (nl_we_op_a), per shot-cell: a step reads prev and curr and writes next
# (12 B; its per-cell factors dt^2 K, 1/rho and the boundary divisor are read once per cell
# for all shots); imaging every sub steps reads the adjoint level and snapshots j-1, j, j+1
# and read-modify-writes the K and 1/rho slabs (32 B); the forward run writes snapshot j
ACOUSTIC_BYTES_FWD = 12.0
ACOUSTIC_BYTES_MODEL = 12.0
ACOUSTIC_BYTES_IMAGING = 32.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c2", "c3"], default="c2",
                    help="c2: 64 isotropic shots on 500x1700 (default); c3: 16 VTI shots on 1000x3000, h = 5 m")
    ap.add_argument("--shots", type=int, default=None, help="shots per batch (default 64 for c2, 16 for c3)")
    ap.add_argument("--nt", type=int, default=None, help="time steps (default 4000 for c2, 6000 for c3)")
    ap.add_argument("--plain-shots", type=int, default=4)
    ap.add_argument("--tile-x", type=int, default=16)
    ap.add_argument("--checkpoint-budget", type=int, default=0, metavar="MIB",
                    help="device memory (MiB) for checkpoints and snapshot slots of the batched form; longer "
                         "checkpoint segments leave fewer steps to recompute (0: the sqrt scheme's minimum)")
    ap.add_argument("--image-fused", type=int, choices=[-1, 0, 1], default=-1,
                    help="1: the imaging condition runs inside the adjoint step kernels (by parts in time); "
                         "0: in a separate imaging kernel after each adjoint step; -1: the engine's default "
                         "(1 on isotropic grids, 0 on VTI grids)")
    ap.add_argument("--engine-preset", action="append", default=[], metavar="KEY=VALUE",
                    help="an engine option set before the autotuner runs, e.g. band_rows=16; repeatable")
    ap.add_argument("--engine-option", action="append", default=[], metavar="KEY=VALUE",
                    help="an engine option (fwi2d_engine_set_option) set after the autotuner ran, e.g. "
                         "shots_per_block_img=8; repeatable")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the small-case gradient error check")
    ap.add_argument("--cpu-steps", type=int, default=1600, help="time steps of the CPU baseline sample")
    ap.add_argument("--summarize", metavar="DIR",
                    help="only summarize the rocprofv3 CSV output under DIR (kernel time per launch shape, "
                         "counters per dispatch) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy 5 times (HBM counter calibration) and exit")
    ap.add_argument("--operator", action="store_true",
                    help="time the operator level instead: nlls_fwi_eco evaluations (forward, residual, "
                         "adjoint) through fwi2d_op / fwi2d_nlls from host arrays, one JSON line")
    ap.add_argument("--fwi", type=int, metavar="ITERS", default=None,
                    help="end-to-end FWI instead: ITERS l-BFGS iterations (regular Wolfe line search) per band on "
                         "the nlls_fwi_eco objective, 48 shots on 600x2000, one JSON line")
    ap.add_argument("--bands", default="5,8,12",
                    help="with --fwi: comma-separated low-pass cut-off frequencies (Hz) of the objective's filter, "
                         "one l-BFGS run per band, low to high")
    ap.add_argument("--fwi-source-hz", type=float, default=10.0,
                    help="with --fwi: peak frequency of the Ricker source of every band (the data are modelled once)")
    ap.add_argument("--fwi-params", choices=["vp", "all"], default="vp",
                    help="with --fwi: invert vp only (gradient mask zero on vs and rho, the default) or all three")
    ap.add_argument("--fwi-mute", type=float, default=0.0, metavar="KM",
                    help="with --fwi: gradient mask zero above this depth, a sin^2 ramp over the next KM (0: none)")
    ap.add_argument("--fwi-a1", type=float, default=0.0,
                    help="with --fwi: regular_wolfe's a1, a first step of a1 |m| / |g| if that is larger than "
                         "a0 / |g| (the reference's default 0)")
    ap.add_argument("--fwi-trials", type=int, default=10, help="with --fwi: trial steps per line search")
    ap.add_argument("--fwi-mask-tapers", action="store_true",
                    help="with --fwi: gradient mask zero in the absorbing taper zones (sides and bottom)")
    ap.add_argument("--acoustic", type=int, metavar="SUB", default=None,
                    help="the acoustic operator (nl_we_op_a) on the C2 geometry: batched gradients with snapshots "
                         "every SUB steps, kept whole")
    ap.add_argument("--step-only", type=int, metavar="EDGE_STREAM", default=None,
                    help="only autotune and time the forward time step of the workload with this edge_stream "
                         "setting (0: every launch of a step on one stream, for per-launch kernel traces), "
                         "print the step time and exit")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_step_iso_64shots_500x1700_traffic.md")
ACOUSTIC_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_astep_acoustic_64shots_500x1700.md")


def measured_traffic(spb=None, path=TRAFFIC_FILE):
    """HBM bytes per forward step from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE per dispatch, calibrated against a copy kernel of known size) for the
    core launch's shots per workgroup spb (the autotuner's choice), or None."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    for key in (["traffic_bytes_per_launch_spb%d:" % spb] if spb else []) + ["traffic_bytes_per_launch:"]:
        for line in lines:
            if line.startswith(key):
                return float(line.split(":")[1])
    return None


def log(msg):
    print("[bench %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


def summarize(d):
    """Per (kernel, grid, workgroup) launch counts and mean durations from rocprofv3's
    kernel_trace.csv, and per-dispatch counter means from counter_collection.csv."""
    import csv
    import glob
    from collections import defaultdict

    def shape(r, key):
        if key in r:
            return r[key]
        return "x".join(r.get(key + "_" + a, "?") for a in "XYZ")

    for path in sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)):
        acc = defaultdict(lambda: [0, 0.0])
        for r in csv.DictReader(open(path)):
            k = (r["Kernel_Name"][:60], shape(r, "Grid_Size"), shape(r, "Workgroup_Size"))
            acc[k][0] += 1
            acc[k][1] += float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
        tot = sum(v[1] for v in acc.values())
        print("# %s: %d dispatches, %.3f s of kernel time" % (os.path.basename(path), sum(v[0] for v in acc.values()),
                                                            tot * 1e-9))
        print("%-60s %-16s %-10s %8s %12s %7s" % ("kernel", "grid", "wg", "calls", "avg_us", "pct"))
        for k, (n, t) in sorted(acc.items(), key=lambda kv: -kv[1][1])[:25]:
            print("%-60s %-16s %-10s %8d %12.2f %6.2f%%" % (k[0], k[1], k[2], n, t / n * 1e-3, 100 * t / tot))
    for path in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
        acc = defaultdict(lambda: [0, 0.0])
        for r in csv.DictReader(open(path)):
            k = (r["Kernel_Name"][:60], shape(r, "Grid_Size"), r["Counter_Name"])
            acc[k][0] += 1
            acc[k][1] += float(r["Counter_Value"])
        print("# %s" % os.path.basename(path))
        print("%-60s %-16s %-12s %8s %16s" % ("kernel", "grid", "counter", "calls", "avg_per_dispatch"))
        for k, (n, v) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
            print("%-60s %-16s %-12s %8d %16.1f" % (k[0], k[1], k[2], n, v / n))


def calibrate_copy():
    import torch
    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))


def converted(m):
    """(vp, vs, rho[, delta, eps]) -> (lambda, mu, rho[, c13, eps]), convert_model
    (we_op.cpp:476-494, 1441-1462)."""
    lam = m[2] * (m[0] ** 2 - 2 * m[1] ** 2)
    mu = m[2] * m[1] ** 2
    out = [lam, mu, m[2]]
    if len(m) == 5:
        out += [np.sqrt(2 * (lam + 2 * mu) * (lam + mu) * m[3] + (lam + mu) ** 2) - mu, m[4]]
    return np.stack(out).astype(np.float32)


WORKLOADS = {
    # name: (nz, nx, h km, dt s, shots, nt, vti, geometry seed, text)
    "c2": (500, 1700, 0.01, 0.001, 64, 4000, False, 2,
           "batched isotropic FWI gradients (forward + checkpointed adjoint): %d shots, 500x1700 grid h=10 m, "
           "nt=%d dt=1 ms, 100-400 receivers/shot, 8-20 Hz Ricker"),
    "c3": (1000, 3000, 0.005, 0.0005, 16, 6000, True, 3,
           "batched VTI FWI gradients (forward + checkpointed adjoint): %d shots, 1000x3000 grid h=5 m, "
           "nt=%d dt=0.5 ms, Thomsen eps in [0, 0.25], delta in [-0.05, 0.15], 8-20 Hz Ricker"),
}


def workload(kind, nshots, nt):
    from fwi2d_amd import workloads as W
    nz, nx, h, dt, _, _, vti, gseed, _ = WORKLOADS[kind]
    true = W.gradient_model(nx, nz, h, h, seed=0)
    if vti:
        true = np.concatenate([true, W.vti_fields(nx, nz, h, h)], 0)
    start = W.smooth_model(true, h, h, 0.3)
    sxz, rxz, freqs = W.config2_geometry(nshots, nx, h, seed=gseed)
    wav = np.stack([W.ricker(nt, dt, f) for f in freqs])
    src = np.stack([wav * np.cos(1.5708), wav * np.sin(1.5708)]).astype(np.float32)
    prob = W.base_problem(nx, nz, h, h, nt, dt, sxz, rxz, vti=int(vti), dt=dt, sub=1)
    return prob, converted(true), converted(start), sxz, rxz, src, dt


class Batch:
    """One engine holding a shot batch, with device-resident traces (torch tensors)."""

    def __init__(self, fwi2d_amd, torch, prob, model_true, model, sxz, rxz, src, nt, dt, tile_x, tuned=True,
                 checkpoint_budget=0, image_fused=-1, presets=()):
        self.torch = torch
        self.eng = fwi2d_amd.Engine(fwi2d_amd.grid_from(prob, nt, dt, 1))
        for kv in presets:
            k, v = kv.split("=", 1)
            self.eng.set_option(k, int(v))
        self.eng.set_option("checkpoint_budget", checkpoint_budget)
        self.eng.set_option("image_fused", image_fused)
        if tuned:
            # batched form: step shapes picked by the autotuner at the first forward run,
            # time loops captured in hipGraphs and replayed
            self.eng.set_option("autotune", 1)
            self.eng.set_option("graph", 1)
        else:
            # plain form: one shot per workgroup, launches issued step by step
            self.eng.set_option("tile_x", tile_x)
            self.eng.set_option("shots_per_block", 1)
        self.eng.set_shots(sxz, rxz)
        self.nt, self.dt = nt, dt
        self.nr = self.eng.nr
        dev = "cuda"
        self.src = torch.from_numpy(np.ascontiguousarray(src)).to(dev)
        self.rcv = torch.zeros((2, self.nr, nt), dtype=torch.float32, device=dev)
        self.adj = torch.zeros_like(self.rcv)
        self.res = torch.zeros_like(self.rcv)
        self.misfit = self.grad_norm = None
        self.grad = torch.zeros((5 if prob.get("vti") else 3, prob["nx"], prob["nz"]), dtype=torch.float32,
                                device=dev)
        # observed data from the true model
        self.eng.set_model(model_true)
        self.eng.forward_device(self.src.data_ptr(), self.rcv.data_ptr(), keep_checkpoints=False)
        self.dobs = self.rcv.clone()
        self.eng.set_model(model)

    def gradient(self):
        t = self.torch
        self.eng.forward_device(self.src.data_ptr(), self.rcv.data_ptr(), keep_checkpoints=True)
        # residual; L2 misfit reduced in FP64 on the device
        t.sub(self.dobs, self.rcv, out=self.res)
        t.cuda.synchronize()
        self.misfit = 0.5 * self.eng.dot_f64(self.res.data_ptr(), 0, 0, self.res.numel())
        # adjoint source of the L2 misfit: residual, time-reversed, sign-flipped
        # (apply_jacobianT's Ht^-1, revert and -alpha applied to Ht r, we_op.cpp:1112-1117)
        self.adj.copy_(t.flip(self.res, dims=[2]))
        t.cuda.synchronize()
        self.eng.adjoint_device(self.adj.data_ptr(), self.grad.data_ptr())
        # gradient norm, FP64 on the device (line-search / solver input)
        self.grad_norm = self.eng.dot_f64(self.grad.data_ptr(), 0, 0, self.grad.numel()) ** 0.5


def grad_error_small(vti=False, fmt=0, image_fused=0):
    """Gradient rel-L2 error vs the FP64 oracle on a reduced single-shot case (largest over
    the lambda / mu / rho components); fmt: checkpoint format (0 FP32, 1 BF16, 2 FP16)."""
    import fwi2d_amd
    from fwi2d_amd import workloads as W
    from oracle.oracle import Oracle, grid_struct
    nx, nz, h, nt, dt = 160, 100, 0.01, 600, 0.001
    true = W.add_ellipses(W.gradient_model(nx, nz, h, h, seed=0), h, h)
    if vti:
        true = np.concatenate([true, W.vti_fields(nx, nz, h, h)], 0)
    start = W.smooth_model(true, h, h, 0.1)
    m0, m1 = converted(true).astype(np.float64), converted(start).astype(np.float64)
    prob = dict(nx=nx, nz=nz, dx=h, dz=h, bc_top=1, bc_bottom=2, bc_left=2, bc_right=2, taper_top=0,
                taper_bottom=20, taper_left=20, taper_right=20, taper_strength=0.05, vti=int(vti))
    sxz = np.array([[0.8, 0.02]])
    rxz = [np.stack([np.arange(2, nx - 2, 4) * h, np.full(len(range(2, nx - 2, 4)), 0.02)], 1)]
    w = W.ricker(nt, dt, 12.0)
    src = np.stack([0 * w[None], w[None]])
    orc = Oracle(64)
    g = grid_struct(dict(prob, nt=nt, dt=dt, sub=1))
    inj = orc.delta_m3(nx, nz, h, h, 0, 0, sxz[:, 0], sxz[:, 1])
    ext = orc.delta_m3(nx, nz, h, h, 0, 0, rxz[0][:, 0], rxz[0][:, 1])
    dobs = orc.propagate(g, m0, inj, src, ext=ext)
    full = np.zeros((1 + nt, 2, nx, nz))
    d = orc.propagate(g, m1, inj, src, ext=ext, full=full)
    adj = -(d - dobs)[:, :, ::-1].copy()
    gref = np.zeros((5 if vti else 3, nx, nz))
    orc.propagate(g, m1, ext, adj, full=full, grad=gref)
    eng = fwi2d_amd.Engine(fwi2d_amd.grid_from(prob, nt, dt, 1))
    eng.set_option("checkpoint_format", fmt)
    eng.set_option("segment", 25)
    eng.set_option("image_fused", image_fused)
    eng.set_model(m1)
    eng.set_shots(sxz, rxz)
    eng.forward(src, keep_checkpoints=True)
    gk = eng.adjoint(adj)
    return max(float(np.linalg.norm(gk[c] - gref[c]) / np.linalg.norm(gref[c])) for c in range(3))


def cpu_baseline(nt_sample=400, kind="c2"):
    """Oracle FP32 restatement (OpenMP, all host cores) on one shot of the same grid for
    nt_sample steps, forward with snapshots + adjoint with imaging; scaled to shots/s."""
    from fwi2d_amd import workloads as W
    from oracle.oracle import Oracle, grid_struct
    orc = Oracle(32)
    nz, nx, h, dt, _, _, vti, gseed, _ = WORKLOADS[kind]
    m = W.gradient_model(nx, nz, h, h, seed=0)
    if vti:
        m = np.concatenate([m, W.vti_fields(nx, nz, h, h)], 0)
    m = converted(m)
    prob = dict(nx=nx, nz=nz, dx=h, dz=h, bc_top=1, bc_bottom=2, bc_left=2, bc_right=2, taper_top=0,
                taper_bottom=30, taper_left=30, taper_right=30, taper_strength=0.05, vti=int(vti))
    sxz, rxz, freqs = W.config2_geometry(1, nx, h, seed=gseed)
    g = grid_struct(dict(prob, nt=nt_sample, dt=dt, sub=1))
    w = W.ricker(nt_sample, dt, float(freqs[0]))
    src = np.stack([0 * w[None], w[None]]).astype(np.float32)
    inj = orc.delta_m3(nx, nz, h, h, 0, 0, sxz[:, 0], sxz[:, 1])
    ext = orc.delta_m3(nx, nz, h, h, 0, 0, rxz[0][:, 0], rxz[0][:, 1])
    full = np.zeros((1 + nt_sample, 2, nx, nz), np.float32)
    grad = np.zeros((5 if vti else 3, nx, nz), np.float32)
    t0 = time.perf_counter()
    d = orc.propagate(g, m, inj, src, ext=ext, full=full)
    orc.propagate(g, m, ext, d[:, :, ::-1].copy(), full=full, grad=grad)
    el = time.perf_counter() - t0
    cells = 2.0 * nt_sample * nx * nz
    return dict(cell_updates_per_s=cells / el, seconds=el, threads=int(orc.lib.orc_num_threads()),
                sample="1 shot, %dx%d grid%s, %d forward + %d adjoint steps (sub=1, imaging every step)"
                % (nz, nx, " (VTI)" if vti else "", nt_sample, nt_sample))


def step_only(args):
    """The forward time step alone: autotuned as in the bench, then 40 eager steps with the
    given edge_stream setting, timed with HIP events on the engine stream."""
    import torch
    import fwi2d_amd
    fwi2d_amd.load()
    wl = WORKLOADS[args.workload]
    ns, nt = args.shots or wl[4], 41
    prob, m_true, m0, sxz, rxz, src, dt = workload(args.workload, ns, nt)
    B = Batch(fwi2d_amd, torch, prob, m_true, m0, sxz, rxz, src, nt, dt, args.tile_x, presets=args.engine_preset)
    B.eng.set_option("graph", 0)
    B.eng.set_option("edge_stream", args.step_only)
    B.eng.set_option("timing", 1)
    for _ in range(2):
        B.eng.forward_device(B.src.data_ptr(), B.rcv.data_ptr(), keep_checkpoints=False)
    us = B.eng.stat("fwd_loop_ns") / 1e3 / B.eng.stat("fwd_loop_steps")
    print(json.dumps({"workload": args.workload, "edge_stream": args.step_only, "step_us": round(us, 2),
                      "presets": args.engine_preset,
                      "tile_x": B.eng.stat("tile_x"), "shots_per_block": B.eng.stat("shots_per_block"),
                      "shots_per_block_edge": B.eng.stat("shots_per_block_edge")}))


def operator_level(args):
    """The reference's call path on the same workload: physical models and host arrays
    through fwi2d_op (analyzeModel, source resampling, device trace maps, H quadrature
    and chain rule) and the nlls_fwi_eco objective (FP64 functional) -- `steps` timed
    evaluations of functional and gradient after `warmup` untimed ones."""
    import fwi2d_amd
    from fwi2d_amd import workloads as W
    fwi2d_amd.load()
    wl = WORKLOADS[args.workload]
    ns = args.shots or wl[4]
    nt = args.nt or wl[5]
    nz, nx, h, dt, _, _, vti, gseed, _ = wl
    true = W.gradient_model(nx, nz, h, h, seed=0)
    if vti:
        true = np.concatenate([true, W.vti_fields(nx, nz, h, h)], 0)
    start = W.smooth_model(true, h, h, 0.3)
    sxz, rxz, freqs = W.config2_geometry(ns, nx, h, seed=gseed)
    wav = np.stack([W.ricker(nt, dt, f) for f in freqs])
    src = np.stack([wav * np.cos(1.5708), wav * np.sin(1.5708)]).astype(np.float32)
    prob = W.base_problem(nx, nz, h, h, nt, dt, sxz, rxz, vti=int(vti), dt=dt, sub=1)
    if args.acoustic is not None:  # nl_we_op_a: (vp, rho), one source and data component
        prob = W.base_problem(nx, nz, h, h, nt, dt, sxz, rxz, dt=dt, sub=args.acoustic)
        prob["acoustic"] = 1
        true, start = true[[0, 2]], start[[0, 2]]
        src = wav[None].astype(np.float32)
    op = fwi2d_amd.WeOp(prob, src)
    dobs = op.apply_forward(true)
    obj = fwi2d_amd.Nlls(op, dobs, normalize=True)
    log("operator and objective ready: %d shots, %dx%d, nt=%d" % (ns, nz, nx, nt))
    for _ in range(args.warmup):
        obj.evaluate(start)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f, g = obj.evaluate(start)
    el = time.perf_counter() - t0
    log("timed evaluations done: %.3f s" % el)
    kind = "acoustic (nl_we_op_a), sub=%d" % args.acoustic if args.acoustic is not None else None
    print(json.dumps({"metric": "operator-level FWI gradient shots/s (nlls_fwi_eco, host arrays)",
                      "value": round(ns * args.steps / el, 4), "unit": "shots/s", "steps": args.steps,
                      "warmup": args.warmup, "s_per_evaluation": round(el / args.steps, 4),
                      "functional_f64": f, "gradient_finite": bool(np.isfinite(g).all()),
                      "config": {"workload": "%s: %d shots, %dx%d, nt=%d%s" % (args.workload, ns, nz, nx, nt,
                                                                             ", VTI" if vti else
                                                                             (", " + kind if kind else ""))}}))


def acoustic_grad_error_small():
    """Acoustic gradient rel-L2 error vs the FP64 oracle on a reduced single-shot case
    (largest over the raw K and 1/rho gradients)."""
    import fwi2d_amd
    from fwi2d_amd import workloads as W
    from oracle.oracle import Oracle, grid_struct
    nx, nz, h, nt, dt, sub = 160, 100, 0.01, 600, 0.001, 2
    phys = W.add_ellipses(W.gradient_model(nx, nz, h, h, seed=0), h, h)
    start = W.smooth_model(phys, h, h, 0.1)
    m0 = np.stack([phys[2] * phys[0] ** 2, 1 / phys[2]]).astype(np.float32).astype(np.float64)
    m1 = np.stack([start[2] * start[0] ** 2, 1 / start[2]]).astype(np.float32).astype(np.float64)
    prob = dict(nx=nx, nz=nz, dx=h, dz=h, bc_top=1, bc_bottom=2, bc_left=2, bc_right=2, taper_top=0,
                taper_bottom=20, taper_left=20, taper_right=20, taper_strength=0.05, vti=0)
    sxz = np.array([[0.8, 0.02]])
    rxz = [np.stack([np.arange(2, nx - 2, 4) * h, np.full(len(range(2, nx - 2, 4)), 0.02)], 1)]
    src = W.ricker(nt, dt, 12.0)[None]
    orc = Oracle(64)
    g = grid_struct(dict(prob, nt=nt, dt=dt, sub=sub))
    inj = orc.delta_m3(nx, nz, h, h, 0, 0, sxz[:, 0], sxz[:, 1])
    ext = orc.delta_m3(nx, nz, h, h, 0, 0, rxz[0][:, 0], rxz[0][:, 1])
    alpha = 0.2508560249 / 1.05
    dobs = orc.propagate_acoustic(g, m0, inj, src, ext=ext, alpha=alpha)
    full = np.zeros((1 + nt // sub, nx, nz))
    d = orc.propagate_acoustic(g, m1, inj, src, ext=ext, full=full, alpha=alpha)
    adj = -(d - dobs)[:, ::-1].copy()
    gref = np.zeros((2, nx, nz))
    orc.propagate_acoustic(g, m1, ext, adj, full=full, grad=gref, alpha=alpha)
    eng = fwi2d_amd.AcousticEngine(fwi2d_amd.grid_from(prob, nt, dt, sub), 1.05)
    eng.set_model(m1)
    eng.set_shots(sxz, rxz)
    eng.forward(src, keep_snapshots=True)
    gk = eng.adjoint(adj)
    eng.close()
    return max(float(np.linalg.norm(gk[c] - gref[c]) / np.linalg.norm(gref[c])) for c in range(2))


def acoustic_cpu_baseline(nt_sample, sub):
    """The oracle's FP32 acoustic propagation (OpenMP over the host cores) on one shot of
    the C2 grid: nt_sample forward steps with snapshots every sub steps + the adjoint run
    with imaging."""
    from fwi2d_amd import workloads as W
    from oracle.oracle import Oracle, grid_struct
    orc = Oracle(32)
    nz, nx, h, dt, _, _, _, gseed, _ = WORKLOADS["c2"]
    phys = W.gradient_model(nx, nz, h, h, seed=0)
    m = np.stack([phys[2] * phys[0] ** 2, 1 / phys[2]]).astype(np.float32)
    prob = dict(nx=nx, nz=nz, dx=h, dz=h, bc_top=1, bc_bottom=2, bc_left=2, bc_right=2, taper_top=0,
                taper_bottom=30, taper_left=30, taper_right=30, taper_strength=0.05, vti=0)
    sxz, rxz, freqs = W.config2_geometry(1, nx, h, seed=gseed)
    g = grid_struct(dict(prob, nt=nt_sample, dt=dt, sub=sub))
    src = W.ricker(nt_sample, dt, float(freqs[0]))[None].astype(np.float32)
    inj = orc.delta_m3(nx, nz, h, h, 0, 0, sxz[:, 0], sxz[:, 1])
    ext = orc.delta_m3(nx, nz, h, h, 0, 0, rxz[0][:, 0], rxz[0][:, 1])
    full = np.zeros((1 + nt_sample // sub, nx, nz), np.float32)
    grad = np.zeros((2, nx, nz), np.float32)
    t0 = time.perf_counter()
    d = orc.propagate_acoustic(g, m, inj, src, ext=ext, full=full)
    orc.propagate_acoustic(g, m, ext, d[:, ::-1].copy(), full=full, grad=grad)
    el = time.perf_counter() - t0
    return dict(cell_updates_per_s=2.0 * nt_sample * nx * nz / el, seconds=el,
                threads=int(orc.lib.orc_num_threads()),
                sample="1 shot, %dx%d grid, %d forward + %d adjoint steps (sub=%d)" % (nz, nx, nt_sample, nt_sample,
                                                                                      sub))


def acoustic_run(args):
    """The acoustic operator (nl_we_op_a) on the C2 geometry: `shots` shots (64) on
    500 x 1700 cells of 10 m, nt = 4000 steps of 1 ms, snapshots every `--acoustic` steps
    kept whole on the device.  `steps` timed batched gradients (forward with snapshots,
    FP64 misfit, adjoint with imaging) after `warmup` untimed ones, all on device
    buffers; then one eager forward run without snapshots timed step by step."""
    import torch
    import fwi2d_amd
    from fwi2d_amd import workloads as W
    fwi2d_amd.load()
    sub = args.acoustic
    nz, nx, h, dt, _, _, _, gseed, _ = WORKLOADS["c2"]
    ns = args.shots or 64
    nt = args.nt or 4000
    phys = W.gradient_model(nx, nz, h, h, seed=0)
    start = W.smooth_model(phys, h, h, 0.3)
    mtrue = np.stack([phys[2] * phys[0] ** 2, 1 / phys[2]]).astype(np.float32)
    m0 = np.stack([start[2] * start[0] ** 2, 1 / start[2]]).astype(np.float32)
    sxz, rxz, freqs = W.config2_geometry(ns, nx, h, seed=gseed)
    src_np = np.stack([W.ricker(nt, dt, f) for f in freqs]).astype(np.float32)
    prob = W.base_problem(nx, nz, h, h, nt, dt, sxz, rxz, dt=dt, sub=sub)
    eng = fwi2d_amd.AcousticEngine(fwi2d_amd.grid_from(prob, nt, dt, sub))
    eng.set_shots(sxz, rxz)
    dev = "cuda"
    src = torch.from_numpy(src_np).to(dev)
    rcv = torch.zeros((eng.nr, nt), dtype=torch.float32, device=dev)
    grad = torch.zeros((2, nx, nz), dtype=torch.float32, device=dev)
    eng.set_model(mtrue)
    eng.forward_device(src.data_ptr(), rcv.data_ptr())
    dobs = rcv.clone()
    eng.set_model(m0)
    log("acoustic workload ready: %d shots, %dx%d, nt=%d, sub=%d" % (ns, nz, nx, nt, sub))
    res = {}

    def gradient():
        eng.forward_device(src.data_ptr(), rcv.data_ptr(), keep_snapshots=True)
        r = dobs - rcv
        res["misfit"] = 0.5 * float(torch.sum(r.double() ** 2))
        adj = torch.flip(r, dims=[1]).contiguous()
        torch.cuda.synchronize()
        eng.adjoint_device(adj.data_ptr(), grad.data_ptr())

    for _ in range(args.warmup):
        gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gradient()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    log("timed acoustic gradients done: %.3f s" % el)
    t1 = time.perf_counter()
    eng.forward_device(src.data_ptr(), rcv.data_ptr(), keep_snapshots=False)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t1
    eng.close()
    shots_per_s = ns * args.steps / el
    step_s = fwd_s / (nt - 1)
    cells = ns * nx * nz
    fwd_bytes = ACOUSTIC_BYTES_FWD * cells + ACOUSTIC_BYTES_MODEL * nx * nz
    grad_bytes = (2 * fwd_bytes * (nt - 1) + (4.0 + ACOUSTIC_BYTES_IMAGING) * cells * (nt // sub + 1))
    err = None if args.no_check else acoustic_grad_error_small()
    cpu = None if args.no_cpu else acoustic_cpu_baseline(800, sub)
    line = {"metric": "acoustic FWI gradient shots/s (nl_we_op_a, batched)", "value": round(shots_per_s, 4),
            "unit": "shots/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1000 * el / args.steps, 2), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "acoustic gradients: %d shots, 500x1700 grid h=10 m, nt=%d dt=1 ms, snapshots "
                                   "every %d steps" % (ns, nt, sub), "shots_per_step": ns, "sub": sub},
            "gcell_updates_per_s": round(shots_per_s * 2.0 * nt * nx * nz / 1e9, 2),
            "misfit_f64": res.get("misfit"),
            "grad_rel_err_vs_fp64": err,
            "grad_err_sample": "1 shot, 160x100 grid, nt=600, sub=2, raw K and 1/rho gradients vs FP64 oracle",
            "roofline": {"bound": "hbm", "kernel": "k_astep + k_ainject per step (eager forward, no snapshots)",
                         "achieved": round(fwd_bytes / step_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(fwd_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": measured_traffic(path=ACOUSTIC_TRAFFIC_FILE) if ns == 64 else None,
                         "avg_step_us": round(step_s * 1e6, 2),
                         "bytes_per_step": fwd_bytes,
                         "gradient_achieved": round(grad_bytes * args.steps / el / 1e9, 1),
                         "gradient_frac": round(grad_bytes * args.steps / el / 1e9 / HBM_PEAK_GBS, 4)}}
    if cpu:
        line["cpu_baseline"] = {"value": round(cpu["cell_updates_per_s"] / (2.0 * nt * nx * nz), 6),
                                "unit": "shots/s", "cores": cpu["threads"], "kind": "port", "sample": cpu["sample"],
                                "gcell_updates_per_s": round(cpu["cell_updates_per_s"] / 1e9, 4)}
    print(json.dumps(line))


def lowpass_filter(fc, dt):
    """A zero-phase low-pass FIR: a Hann-windowed sinc of cut-off fc (Hz), 1.5 periods of
    the cut-off on each side, unit gain at zero frequency."""
    half = int(np.ceil(1.5 / (fc * dt)))
    k = np.arange(-half, half + 1)
    f = 2 * fc * dt * np.sinc(2 * fc * dt * k) * np.hanning(2 * half + 3)[1:-1]
    return (f / f.sum()).astype(np.float32)


def fwi_run(args):
    """The end-to-end FWI workload C4 (48 shots, 600x2000, 4000 samples of 1 ms): data
    modelled once from the true model with a broadband Ricker source, then, band by band
    (low-pass cut-offs from --bands, low to high), `--fwi` l-BFGS iterations (regular
    Wolfe line search at the reference defaults unless --fwi-a1) on nlls_fwi_eco with the
    band's filter (the reference's filter= option: observed data and synthetics convolved
    with the same zero-phase low-pass), each band starting from the previous band's model
    (the first from the 500 m smoothed true model).  One progress line per objective
    evaluation, one JSON line at the end with the functional history and the model
    errors per band."""
    import fwi2d_amd
    from fwi2d_amd import workloads as W
    fwi2d_amd.load()
    bands = [float(b) for b in args.bands.split(",")]
    prob, true, start, src = W.config4(nshots=args.shots or 48, f0=args.fwi_source_hz)
    ns = len(prob["sxz"])
    dt = prob["mp"]["dt_data"]
    h = prob["dz"]
    depth = np.arange(prob["nz"]) * h
    deep = depth >= 0.2  # below the near-surface zone of the sources and receivers (20 m)
    err = lambda x: float(np.linalg.norm(x - true) / np.linalg.norm(true))
    vp_err = lambda x: float(np.linalg.norm(x[0] - true[0]) / np.linalg.norm(true[0]))
    vp_err_deep = lambda x: float(np.linalg.norm((x[0] - true[0])[:, deep]) / np.linalg.norm(true[0][:, deep]))
    gmask = np.ones_like(true, dtype=np.float32)
    if args.fwi_mute > 0:
        ramp = np.clip((depth - args.fwi_mute) / args.fwi_mute, 0.0, 1.0)
        gmask[:] = (np.sin(0.5 * np.pi * ramp) ** 2)[None, None, :]
    if args.fwi_params == "vp":
        gmask[1:] = 0.0
    nx, nz = prob["nx"], prob["nz"]
    inner = np.zeros((nx, nz), bool)  # outside the absorbing taper zones
    inner[prob["taper_left"]:nx - prob["taper_right"], prob["taper_top"]:nz - prob["taper_bottom"]] = True
    if args.fwi_mask_tapers:
        gmask[:, ~inner] = 0.0
    vp_err_inner = lambda x: float(np.linalg.norm((x[0] - true[0])[inner]) / np.linalg.norm(true[0][inner]))
    op = fwi2d_amd.WeOp(prob, src)
    dobs = op.apply_forward(true)
    log("observed data modelled: %d shots, %d traces" % (ns, 2 * op.nr))
    m = start
    out = []
    t_all = time.perf_counter()
    for fc in bands:
        filt = lowpass_filter(fc, dt)
        t1 = time.perf_counter()
        fwi2d_amd.filter_apply(filt, dobs)  # one pass over every trace: an evaluation makes two
        t_filter = time.perf_counter() - t1
        obj = fwi2d_amd.Nlls(op, dobs, gmask=gmask, normalize=True, filter=filt)
        count = [0]

        def evaluate(x):
            f, g = obj.evaluate(x)
            count[0] += 1
            log("band %g Hz evaluation %d: functional %.6e" % (fc, count[0], f))
            return f, g
        t0 = time.perf_counter()
        m1, hist, k, info = fwi2d_amd.lbfgs(evaluate, m, niter=args.fwi, max_trial=args.fwi_trials, a1=args.fwi_a1,
                                            report=True)
        el = time.perf_counter() - t0
        dv = m1[0] - m[0]
        ix, iz = np.unravel_index(int(np.argmax(np.abs(dv))), dv.shape)
        out.append({"band_hz": fc, "filter_taps": int(filt.size), "filter_pass_seconds": round(t_filter, 3),
                    "iterations": k, "evaluations": count[0],
                    "seconds": round(el, 1), "functional_history": [float(v) for v in hist],
                    "model_rel_err": [err(m), err(m1)], "vp_rel_err": [vp_err(m), vp_err(m1)],
                    "vp_rel_err_below_200m": [vp_err_deep(m), vp_err_deep(m1)],
                    "vp_rel_err_outside_tapers": [vp_err_inner(m), vp_err_inner(m1)],
                    "vp_range": [float(m1[0].min()), float(m1[0].max())],
                    "max_vp_change": [float(dv[ix, iz]), round(ix * h, 3), round(iz * h, 3)],
                    "trials": [[int(r[0]), int(r[1]), float(r[2]), float(r[3])] for r in info["trials"]]})
        log("band %g Hz: %d iterations, %d evaluations, %.1f s, functional %.4e -> %.4e, vp error %.4f -> %.4f "
            "(outside tapers %.4f -> %.4f), vp in [%.3f, %.3f], largest change %.3f at x %.2f z %.2f km"
            % (fc, k, count[0], el, hist[0], hist[k], vp_err(m), vp_err(m1), vp_err_inner(m), vp_err_inner(m1),
               m1[0].min(), m1[0].max(), dv[ix, iz], ix * h, iz * h))
        m = m1
        obj.close()
    op.close()
    print(json.dumps({"metric": "end-to-end FWI (l-BFGS, nlls_fwi_eco with filter= bands)", "bands": out,
                      "seconds": round(time.perf_counter() - t_all, 1),
                      "model_rel_err": [err(start), err(m)], "vp_rel_err": [vp_err(start), vp_err(m)],
                      "vp_rel_err_below_200m": [vp_err_deep(start), vp_err_deep(m)],
                      "config": {"workload": "%d shots, 600x2000 h=10 m, 4000 samples of 1 ms, %g Hz Ricker, "
                                             "low-pass bands %s Hz, %d l-BFGS iterations per band, inverting %s, "
                                             "mute %g km, a1 %g, tapers %s"
                                             % (ns, args.fwi_source_hz, args.bands, args.fwi,
                                                "vp" if args.fwi_params == "vp" else "vp, vs, rho", args.fwi_mute,
                                                args.fwi_a1, "masked" if args.fwi_mask_tapers else "free")}}))


def main():
    args = parse()
    if args.summarize:
        summarize(args.summarize)
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.step_only is not None:
        step_only(args)
        return
    if args.operator:  # with --acoustic SUB: the acoustic operator at the operator level
        operator_level(args)
        return
    if args.fwi is not None:
        fwi_run(args)
        return
    if args.acoustic is not None:
        acoustic_run(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine batches the shots of one MI355X; across GPUs "
              "the shots are split, one engine per GPU, and only the raw gradient is all-reduced.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import fwi2d_amd
    fwi2d_amd.load()
    wl = WORKLOADS[args.workload]
    ns = args.shots or wl[4]
    nt = args.nt or wl[5]
    vti = wl[6]
    bytes_per_cell = 44.0 if vti else BYTES_PER_CELL_UPDATE_FWD
    prob, m_true, m0, sxz, rxz, src, dt = workload(args.workload, ns, nt)
    nx, nz = prob["nx"], prob["nz"]
    cells_per_grad = 2.0 * nt * nx * nz  # forward + adjoint, per shot

    # ---- batched form -------------------------------------------------------
    log("workload ready: %d shots, %dx%d, nt=%d" % (ns, nz, nx, nt))
    B = Batch(fwi2d_amd, torch, prob, m_true, m0, sxz, rxz, src, nt, dt, args.tile_x,
              checkpoint_budget=args.checkpoint_budget, image_fused=args.image_fused, presets=args.engine_preset)
    log("observed data modelled")
    for kv in args.engine_option:
        k, v = kv.split("=", 1)
        B.eng.set_option(k, int(v))
    for _ in range(args.warmup):
        B.gradient()
    torch.cuda.synchronize()
    log("warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        B.gradient()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    log("timed gradients done: %.3f s" % el)
    ms_per_step = 1000.0 * el / args.steps
    shots_per_s = ns * args.steps / el
    gcells = shots_per_s * cells_per_grad / 1e9
    grad_norm, misfit = B.grad_norm, B.misfit
    grad_gbs = BYTES_PER_GRADIENT_CELL_STEP[vti] * ns * nt * nx * nz / (ms_per_step / 1000.0) / 1e9
    traffic = measured_traffic(B.eng.stat("shots_per_block")) if (args.workload, ns) == ("c2", 64) else None

    # ---- dominant kernel: fused step, timed with HIP events on its stream ----
    # (eager launches: one step = the core-box launch plus the edge-band launches beside it)
    image_fused = B.eng.stat("image_fused")
    tune = {"tile_x": B.eng.stat("tile_x"), "shots_per_block": B.eng.stat("shots_per_block"),
            "shots_per_block_edge": B.eng.stat("shots_per_block_edge"), "edge_stream": B.eng.stat("edge_stream"),
            "report": B.eng.tune_report()}
    graph_replays = B.eng.stat("graph_replays")
    recompute = B.eng.stat("recompute_steps")
    B.eng.set_option("graph", 0)
    B.eng.set_option("timing", 1)
    B.eng.forward_device(B.src.data_ptr(), B.rcv.data_ptr(), keep_checkpoints=False)
    loop_ns = B.eng.stat("fwd_loop_ns")
    nsteps = B.eng.stat("fwd_loop_steps")
    step_s = loop_ns * 1e-9 / nsteps
    # one more gradient, eager, for the time split of its adjoint run: the recomputation
    # steps at the forward step's time, the rest the adjoint steps with their imaging
    rec0 = B.eng.stat("recompute_steps")
    B.gradient()
    adj_s = B.eng.stat("adj_ns") * 1e-9
    rec_steps = B.eng.stat("recompute_steps") - rec0
    B.eng.set_option("timing", 0)
    adjoint_step_us = 1e6 * (adj_s - rec_steps * step_s) / (nt - 1)
    cells_per_launch = ns * nx * nz
    achieved = bytes_per_cell * cells_per_launch / step_s / 1e9
    segment = B.eng.stat("segment")
    del B
    torch.cuda.empty_cache()

    # ---- plain form: one shot per launch, host-driven steps -----------------
    P = args.plain_shots
    t_plain = 0.0
    plain_shots_per_s = None
    for s in range(P):
        one = Batch(fwi2d_amd, torch, prob, m_true, m0, sxz[s:s + 1], rxz[s:s + 1], src[:, s:s + 1], nt, dt,
                    args.tile_x, tuned=False)
        one.gradient()  # warm
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        one.gradient()
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t1
        del one
    if P > 0:
        plain_shots_per_s = P / t_plain
        log("plain form done")

    grad_err = None if args.no_check else grad_error_small(vti, 0, args.image_fused)
    grad_err16 = None
    if vti and not args.no_check:
        grad_err16 = {"bf16": grad_error_small(True, 1, args.image_fused),
                      "fp16": grad_error_small(True, 2, args.image_fused)}
    log("gradient check done")
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_steps if not vti else max(args.cpu_steps // 4, 50),
                                                  args.workload)
    log("cpu baseline done")

    line = {
        "metric": METRIC,
        "value": round(shots_per_s, 4),
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl[8] % (ns, nt),
                   "shots_per_step": ns, "nx": nx, "nz": nz, "nt": nt, "autotuned": tune,
                   "hipgraph_replays": graph_replays},
        "gcell_updates_per_s": round(gcells, 2),
        "plain_one_shot_per_launch_shots_per_s": plain_shots_per_s and round(plain_shots_per_s, 4),
        "batched_vs_plain": plain_shots_per_s and round(shots_per_s / plain_shots_per_s, 3),
        "recompute_steps_per_gradient": int(recompute // max(args.steps + args.warmup, 1)),
        "time_split_eager_s": {"forward": round(step_s * (nt - 1), 3), "adjoint_run": round(adj_s, 3),
                               "recompute_steps": int(rec_steps),
                               "adjoint_step_us_with_imaging": round(adjoint_step_us, 1)},
        "checkpoint_segment": segment,
        "checkpoint_budget_mib": args.checkpoint_budget,
        "image_fused": image_fused,
        "engine_options": args.engine_preset + args.engine_option,
        "grad_rel_err_vs_fp64": grad_err,
        "grad_err_sample": "1 shot, 160x100 grid%s, nt=600, raw lambda/mu/rho gradients vs FP64 oracle"
                           % (" (VTI)" if vti else ""),
        "grad_norm_f64": grad_norm,
        "misfit_f64": misfit,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_step (fused time step of all shots: core-box launch + edge-band launches)",
                     "avg_launch_us": round(step_s * 1e6, 2),
                     "bytes_per_cell_update": bytes_per_cell, "cells_per_launch": cells_per_launch,
                     # the whole gradient (forward + checkpointed adjoint with imaging) against the
                     # same peak: algorithmic bytes of one gradient of the batch / its wall time
                     "gradient_achieved": round(grad_gbs, 1), "gradient_frac": round(grad_gbs / HBM_PEAK_GBS, 4),
                     "gradient_bytes_per_cell_step": BYTES_PER_GRADIENT_CELL_STEP[vti]},
    }
    if grad_err16:
        line["grad_rel_err_vs_fp64_16bit_checkpoints"] = grad_err16
    if cpu:
        line["cpu_baseline"] = {"value": round(cpu["cell_updates_per_s"] / cells_per_grad, 6), "unit": "shots/s",
                                "cores": cpu["threads"], "kind": "port", "sample": cpu["sample"],
                                "gcell_updates_per_s": round(cpu["cell_updates_per_s"] / 1e9, 4)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
