"""Benchmark of the pSi3D time step on one MI355X.

Workload (default): the synthetic stratified elliptic lake of 2048 x 2048
cells, dx = 50 m, 60 stretched layers (0.5 -> 5 m), GRF bathymetry, tanh
thermocline, sinusoidal wind from 30 deg, TVD transport + two-equation
turbulence, dt = 30 s (`--nx 1024 --ny 512` gives the 1024 x 512, dx = 100 m,
dt = 60 s lake; both have the first internal mode's Courant number c1 dt / dx
near 0.35).  A step is one full si3d.f90 time step (leapfrog fd + settrap + 2
trapezoidal fd + settrap2 + save) with all three free-surface solves to a
1e-10 FP64 relative true residual.  Inputs are resident in HBM before the
timed region.

`value` is the plain FP64 path.  The learned legs: the network is trained in a
context of its own on pairs the plain solver records under random winds, then
its weights drive the initial guess of FP64 and mixed-precision runs of the
measured lake under its own (unseen) wind, from the same initial state and
over the same steps as the plain leg; each reports its throughput, the speed-up
over plain FP64, iterations and its free surface's difference from plain FP64.

Prints one JSON line: wet-cell-steps/s (value), PCG iterations per step, the
roofline of the dominant kernel (HIP events around every launch in the timed
region), the learned legs and a CPU baseline (the FP64 oracle on the host, a
bounded sample).
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

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per wet cell for one launch of each kernel: distinct state
# arrays read plus distinct arrays written that a later phase consumes, counted
# once per wet cell (phase-private scratch, index arrays and 2-D arrays not
# counted), as (arrays of the storage type, FP64 arrays).  Every 3-D field,
# the turbulence state included, has the storage type (FP64, or FP32 in mixed
# mode), so the second count is zero.
ARRAYS_PER_CELL = {
    "exmom_x": (10, 0), "exmom_y": (10, 0),      # uhp vhp wp up upp uhpp hup hupp hp -> ex
    "matmom_x": (8, 0), "matmom_y": (8, 0),      # ex Av rhop hup upp (uhpp at the bed) -> cfx agx arx
    "matcon": (2, 0),                            # uhpp vhpp
    "slayer_h": (6, 0),                          # hp hup hvp -> h hu hv
    "vel": (8, 0),                               # cfx agx arx cfy agy ary (hu hv at the top) -> uh vh
    "continuity1": (5, 0),                       # uh uhpp vh vhpp -> wp
    "exsal_hflux": (10, 0),                      # salpp uhpp uh vhpp vh hup hvp hpp -> fx_ew fx_ns
    "exsal_imsal": (11, 0),                      # fx_ew fx_ns salpp hpp hp hupp hvpp wp h Dv -> sal
    "exsal_rhs": (10, 0),                        # fx_ew fx_ns salpp hpp hp hupp hvpp wp -> bex
    "imsal": (5, 0),                             # bex Dv h (hpp, wp at the top) -> sal
    "vel2": (3, 0),                              # -> h hu hv
    "turbulence": (14, 0),                       # hp rhop up vp Av Dv q2p q2pp q2lp q2lpp -> Av Dv q2 q2l
    "turb_cell": (10, 0),                        # hp rhop up vp Av Dv q2p q2lp -> Av Dv
    "settrap_cell": (36, 0),                     # 15 read, 21 written (turbulence state included)
    "settrap2_cell": (21, 0),                    # 10 read, 11 written
    "save_cell": (21, 0),                        # trapezoidal save: 8 read, 13 written
    "continuity2": (3, 0),                       # uhp vhp -> wp
}


# HBM bytes per wet cell per launch measured with rocprofv3 FETCH_SIZE /
# WRITE_SIZE (scaled by a 512 MiB device copy) on the default workload
# (2048 x 2048 x 60, dx = 50 m, dt = 30 s, FP64):
# profiles/hbm_traffic_ellipse_2048x2048x60_dx50_fp64.md.  Reported as
# roofline.traffic for that workload only.
TRAFFIC_FP64_2048x2048x60_DX50 = {
    "exmom_x": 160.9, "exmom_y": 137.0, "matmom_x": 89.6, "matmom_y": 78.5, "matcon": 27.6, "slayer_h": 52.2,
    "vel": 73.8, "continuity1": 73.6, "exsal_hflux": 117.8, "exsal_rhs": 116.9, "imsal": 65.7, "vel2": 27.6,
    "turbulence": 220.7, "turb_cell": 140.4, "settrap_cell": 304.7, "settrap2_cell": 180.2, "save_cell": 178.0,
    "continuity2": 40.7,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--layers", type=int, default=60)
    ap.add_argument("--dx", type=float, default=None,
                    help="grid spacing (m); default 50 for the 2048 x 2048 lake, 100 otherwise")
    ap.add_argument("--dt", type=float, default=None,
                    help="time step (s); default 60 at dx = 100 m and 30 at dx = 50 m, which keeps the "
                         "first internal mode's Courant number c1 dt / dx at about 0.35 (dt = 60 s at 50 m doubles "
                         "it and the scheme's checkerboard internal-wave mode grows: the oracle and the GPU both "
                         "diverge within 20 steps)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--precond", default="mg", choices=["jacobi", "line", "mg"])
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-procs", type=int, default=min(16, os.cpu_count() or 1),
                    help="host cores for the CPU baseline: one single-threaded copy of the oracle per core, timed "
                         "together (1: one copy)")
    ap.add_argument("--cpu-side", type=int, default=512,
                    help="for lakes above 1024 x 512 cells the CPU baseline runs the same generator on a "
                         "CPU-SIDE x CPU-SIDE lake of the same spacing (the full lake needs minutes per step)")
    ap.add_argument("--no-timers", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-initial-guess legs")
    ap.add_argument("--train-steps", type=int, default=300, help="optimiser steps of the network (batch 4)")
    ap.add_argument("--train-pairs", type=int, default=192,
                    help="(right-hand side, solution) pairs recorded from plain solves under random winds")
    ap.add_argument("--train-wind", default="0.005,0.05", metavar="LO,HI",
                    help="range (Pa) of the random wind stresses of the training run (a new wind every 10 min)")
    ap.add_argument("--flex-period", type=int, default=0,
                    help="also run the learned correction stage inside flexible CG every N iterations")
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="advance N lakes of uneven size (sides 128-768, 20-80 layers, random wind) together")
    ap.add_argument("--side-max", type=int, default=768)
    ap.add_argument("--tracers", type=int, default=0, metavar="N",
                    help="advance N conservative tracers as well (step profiles)")
    ap.add_argument("--surfbc", type=int, default=20, choices=[1, 2, 3, 20],
                    help="surface forcing: 20 wind series (default), 1-3 heat budget over the same wind")
    ap.add_argument("--river", type=float, default=0.0, metavar="Q",
                    help="a channel of the same size with flow Q (m3/s) entering at the west end and leaving at "
                         "the east end (open boundaries of type 2) instead of the elliptic lake")
    ap.add_argument("--mixed-error", type=int, default=0, metavar="STEPS",
                    help="advance the lake in FP64 and in mixed precision side by side for STEPS steps and report "
                         "the per-field max / RMS difference (one JSON line) instead of the throughput")
    ap.add_argument("--tol-floor", type=int, default=0, metavar="STEPS",
                    help="like --mixed-error, but the second run is FP64 with the free surface solved to 1e-12: "
                         "how far FP64 itself drifts with the solver tolerance")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (calibrates HBM byte counters) and exit")
    ap.add_argument("--unet-check", type=int, default=0, metavar="N",
                    help="only run the network's forward and backward pass three times on a random N x N input "
                         "(s3g_unet_check; a workload for profiling the network's kernels) and exit")
    return ap.parse_args()


def cpu_baseline(deck, nsteps):
    """The FP64 oracle (C++ restatement of the reference, one core) on the same
    deck: init and step 1 (two fd calls instead of three) excluded, steps
    2 .. nsteps+1 timed."""
    from oracle.oracle import Oracle, build
    build()
    o = Oracle(deck)
    o.step(1)
    t0 = time.perf_counter()
    for n in range(2, nsteps + 2):
        o.step(n)
    dt = time.perf_counter() - t0
    return dt, o


def _cpu_worker(deck, nsteps, barrier, out):
    from oracle.oracle import Oracle
    o = Oracle(deck)
    o.step(1)
    barrier.wait()
    t0 = time.perf_counter()
    for n in range(2, nsteps + 2):
        o.step(n)
    out.put(time.perf_counter() - t0)


def cpu_baseline_procs(deck, nsteps, procs):
    """The oracle on `procs` host cores: one copy of the lake per core, each a
    single-threaded process (started fresh, not forked from this one), set up
    and stepped once, then steps 2 .. nsteps+1 timed together after a common
    barrier.  Returns the longest of the timed spans."""
    import multiprocessing as mp
    import queue
    from oracle.oracle import build
    build()
    ctx = mp.get_context("spawn")
    barrier, out = ctx.Barrier(procs, timeout=900), ctx.Queue()
    ps = [ctx.Process(target=_cpu_worker, args=(deck, nsteps, barrier, out)) for _ in range(procs)]
    for p in ps:
        p.start()
    spans = []
    try:
        while len(spans) < procs:
            try:
                spans.append(out.get(timeout=5))
            except queue.Empty:
                if any(p.exitcode not in (None, 0) for p in ps):
                    raise RuntimeError("a CPU-baseline process failed")
    finally:
        for p in ps:
            p.join(timeout=10)
            if p.is_alive():
                p.terminate()
    return max(spans)


def copy_probe(mib):
    """A plain device-to-device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of the profiled kernels."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4}))


def unet_probe(n):
    """Forward + backward of the network on a random n x n input, three times
    (the first call includes the allocation), for kernel profiles."""
    from si3d_amd.model import unet_check, unet_num_weights
    rng = np.random.default_rng(3)
    wv = (0.05 * rng.standard_normal(unet_num_weights())).astype(np.float32)
    x3 = rng.normal(size=(3, n, n)).astype(np.float32)
    target = rng.normal(size=(n, n)).astype(np.float32)
    mask = (rng.random((n, n)) > 0.2).astype(np.float32)
    for rep in range(3):
        t0 = time.perf_counter()
        _, _, loss = unet_check(wv, x3, target, mask)
        print(json.dumps({"unet_check": f"{n}x{n}", "rep": rep, "seconds": round(time.perf_counter() - t0, 3),
                          "loss": loss}), flush=True)


def stream_bandwidth(mib=2048, reps=10):
    """Read + write GB/s of the library's own streaming-copy kernel
    (s3g_probe_stream: 16-byte vector loads and stores) over `mib` MiB."""
    import ctypes
    from si3d_amd import _lib
    gbs = ctypes.c_double(0.0)
    if _lib.load().s3g_probe_stream(ctypes.c_size_t(mib << 20), reps, ctypes.byref(gbs)) != 0:
        return None
    return round(gbs.value, 1)


def copy_bandwidth(mib=1024, reps=10):
    """Achievable HBM bandwidth on this device: a device-to-device copy of `mib`
    MiB (read + write bytes / time), timed with HIP events, through the HIP
    runtime the library itself uses (a second runtime in the process, such as
    torch's, would not see the device)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    ok = (hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
          and hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
          and hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0)
    ms = ctypes.c_float(0.0)
    if ok:
        hip.hipMemset(src, 0, ctypes.c_size_t(nbytes))
        hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(dst, src, ctypes.c_size_t(nbytes), 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    for p_ in (src, dst):
        if p_.value:
            hip.hipFree(p_)
    for e in (e0, e1):
        if e.value:
            hip.hipEventDestroy(e)
    if not ok or ms.value <= 0:
        return None
    return 2 * nbytes * reps / (ms.value * 1e-3) / 1e9


def mixed_error(a, floor=False):
    """The mixed-precision run against the FP64 run of the same lake: per-field
    max and RMS difference over the wet cells after a.mixed_error steps (and at
    every quarter of the run), absolute and relative to the FP64 field's max /
    RMS.  Both runs solve the free surface to the same tolerance.  With
    `floor` the second run is FP64 solved to 1e-12 instead (a.tol_floor
    steps): the spread FP64 itself has at the bench's tolerance."""
    from si3d_amd import lake
    from si3d_amd.model import GpuLake
    nsteps = a.tol_floor if floor else a.mixed_error
    deck = lake.ellipse(nx=a.nx, ny=a.ny, dx=a.dx, nlayers=a.layers, dt=a.dt, nts=nsteps + 2, seed=1)
    f64 = GpuLake(deck, precision="fp64", preconditioner=a.precond, tol=a.tol)
    mix = (GpuLake(deck, precision="fp64", preconditioner=a.precond, tol=1e-12) if floor else
           GpuLake(deck, precision="mixed", preconditioner=a.precond, tol=a.tol))
    second = "fp64_tol1e-12" if floor else "mixed"
    fields = ["s", "uh", "vh", "wp", "sal", "h", "q2", "Av"]
    h0 = f64.get("h")
    wet3 = h0 > 0.0
    lm = f64.lm

    def errors():
        out = {}
        wet_now = f64.get("h") > 0.0
        for name in fields:
            x, y = f64.get(name), mix.get(name)
            if x.ndim == 1:
                x, y = x[:lm], y[:lm]
            else:
                m = wet_now[:lm] | wet3[:lm]
                x, y = x[:lm][m], y[:lm][m]
            d = np.abs(y.astype(np.float64) - x.astype(np.float64))
            ax = np.abs(x.astype(np.float64))
            rms_x = float(np.sqrt(np.mean(x.astype(np.float64) ** 2)))
            rms_d = float(np.sqrt(np.mean(d ** 2)))
            out[name] = {"max": float(d.max()), "rms": rms_d, "max_rel": float(d.max() / max(ax.max(), 1e-300)),
                         "rms_rel": rms_d / max(rms_x, 1e-300)}
        return out

    checkpoints = {}
    t0 = time.perf_counter()
    for n in range(1, nsteps + 1):
        f64.step(n)
        mix.step(n)
        if n % max(1, nsteps // 4) == 0 or n == nsteps:
            checkpoints[n] = errors()
            print(f"{second} vs fp64: step {n} of {nsteps}, {time.perf_counter() - t0:.0f} s, "
                  f"s max_rel {checkpoints[n]['s']['max_rel']:.2e}", file=sys.stderr, flush=True)
    print(json.dumps({"report": (f"FP64 with the free surface solved to 1e-12 vs FP64 at {a.tol:g}, same lake" if floor
                                 else "mixed precision vs FP64, same lake, same free-surface tolerance"),
                      "workload": f"ellipse lake {a.nx}x{a.ny}x{a.layers}, dx={a.dx:g} m, dt={a.dt:g} s, {nsteps} steps, "
                                  f"PCG ({a.precond}) tol {a.tol:g}",
                      "errors_after": {str(k): v for k, v in checkpoints.items()},
                      "pcg_iters_per_step": {"fp64": f64.stats()["pcg_iters_total"] / nsteps,
                                             second: mix.stats()["pcg_iters_total"] / nsteps}}))


def per_solve_iterations(g, deck, n_first, nsteps):
    """PCG iterations of each free-surface solve of a step (the leapfrog
    solve, then the trapezoidal iterations), averaged over nsteps steps run
    call by call as the time loop does (si3d.f90:324-393), untimed."""
    niter = max(1, deck.get("niter", 2))
    rows = []
    for n in range(n_first, n_first + nsteps):
        thrs = n * deck["dt"] / 3600.0
        row = []
        g.fd(n, 1, lastiter=0, thrs=thrs)
        row.append(g.stats()["pcg_iters_last"])
        g.settrap()
        for it in range(1, niter + 1):
            g.fd(n, 2, lastiter=1 if it == niter else 0, thrs=thrs, iter=it)
            row.append(g.stats()["pcg_iters_last"])
            if it < niter:
                g.settrap2()
        g.save(2)
        rows.append(row)
    return [round(float(v), 2) for v in np.mean(np.asarray(rows, dtype=np.float64), axis=0)]


def train_on_random_winds(a, deck):
    """The network's training run: the same lake under a sequence of random
    winds (seeded; a new stress and direction every 10 min, amplitudes in
    --train-wind) in a context of its own, plain FP64 PCG recording
    --train-pairs solves, then --train-steps optimiser steps.  Returns the
    weights and a summary; the context is released before the measured legs."""
    from si3d_amd import lake
    from si3d_amd.model import GpuLake
    lo, hi = (float(v) for v in a.train_wind.split(","))
    nrec = (a.train_pairs + 2) // 3 + 1  # 2 solves on step 1, 3 on each later step
    td = dict(deck)
    td["nts"] = max(deck["nts"], nrec + 2)
    td["surfbc"] = lake.random_wind_series(td["nts"] * deck["dt"], 600.0, amp=(lo, hi), seed=11)
    t0 = time.perf_counter()
    g = GpuLake(td, precision="fp64", preconditioner=a.precond, tol=a.tol)
    g.learn_setup(capacity=a.train_pairs, seed=1)
    for n in range(1, nrec + 1):
        g.step(n)
    g.synchronize()
    t_rec = time.perf_counter() - t0
    g.learn_record(False)
    t1 = time.perf_counter()
    loss0 = g.learn_train(1, batch=4, lr=0.0)
    loss1 = g.learn_train(a.train_steps, batch=4, lr=1e-3)
    t_tr = time.perf_counter() - t1
    info = g.learn_info()
    w = g.learn_weights()
    g.close()
    print(f"bench: trained on {info['pairs']} pairs ({nrec} steps under random winds) in {t_tr:.1f} s, "
          f"loss {loss0:.3e} -> {loss1:.3e}", file=sys.stderr, flush=True)
    return w, {"pairs": info["pairs"], "record_steps": nrec, "forcing": f"random wind stress {lo:g}-{hi:g} Pa, any "
               f"direction, a new wind every 600 s (seed 11)", "train_steps": info["train_steps"],
               "batch": 4, "record_s": round(t_rec, 2), "train_s": round(t_tr, 2), "loss_before": loss0,
               "loss_after": loss1,
               "net": f"U-Net 4 levels, 32 base channels, bf16 MFMA, {info['height']}x{info['width']}"}


def run_learned(a, deck, n_wet, weights, precision, s_plain, plain_value, flex=0):
    """A measured leg with the learned initial guess (weights trained under
    other forcings) from the same initial state as the plain FP64 leg, over
    the same warmup + timed steps: throughput, iterations, and the free
    surface's difference from the plain FP64 leg after those steps."""
    from si3d_amd.model import GpuLake
    g = GpuLake(deck, precision=precision, preconditioner=a.precond, tol=a.tol)
    g.learn_setup(capacity=0, seed=1)
    g.learn_weights(weights)
    g.learn_record(False)
    g.learn_mode(1)
    if flex:
        g.learn_flexible(flex)
    for n in range(1, a.warmup + 1):
        g.step(n)
    g.synchronize()
    st0 = g.stats()
    t0 = time.perf_counter()
    for n in range(a.warmup + 1, a.warmup + a.steps + 1):
        g.step(n)
    g.synchronize()
    el = time.perf_counter() - t0
    st1 = g.stats()
    s = g.get("s")[:g.lm]
    info = g.learn_info()
    per_solve = per_solve_iterations(g, deck, a.warmup + a.steps + 1, 3) if deck.get("itrap", 1) == 1 else None
    g.close()
    value = n_wet * a.steps / el
    d = float(np.max(np.abs(s - s_plain)) / max(float(np.max(np.abs(s_plain))), 1e-300))
    print(f"bench: learned {precision}{' flex ' + str(flex) if flex else ''}: {1e3 * el / a.steps:.2f} ms/step, "
          f"{(st1['pcg_iters_total'] - st0['pcg_iters_total']) / a.steps:.2f} iters/step", file=sys.stderr,
          flush=True)
    return {"precision": precision, "value": round(value, 1), "unit": "wet-cell-steps/s",
            "ms_per_step": round(1e3 * el / a.steps, 4), "speedup_vs_plain_fp64": round(value / plain_value, 4),
            "pcg_iters_per_step": round((st1["pcg_iters_total"] - st0["pcg_iters_total"]) / a.steps, 2),
            "pcg_iters_per_solve": per_solve, "pcg_relres_last": st1["pcg_relres_last"],
            "guess_error_ratio": round(info["guess_error_ratio"], 4), "guess_solves": info["guess_solves"],
            "guess_rejected": info["guess_rejected"], "flex_period": flex, "corrections": info["corrections"],
            "zeta_max_rel_diff_vs_plain_fp64": d}


def main():
    a = parse()
    if a.dx is None:
        a.dx = 50.0 if (a.nx, a.ny) == (2048, 2048) else 100.0
    if a.dt is None:
        a.dt = 60.0 * min(1.0, a.dx / 100.0)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.unet_check:
        unet_probe(a.unet_check)
        return
    if a.mixed_error or a.tol_floor:
        mixed_error(a, floor=bool(a.tol_floor))
        return
    if a.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the multi-GPU decomposition exists as a design, not as code)", file=sys.stderr)
        sys.exit(2)
    from si3d_amd import lake
    from si3d_amd.model import GpuEnsemble, GpuLake

    nts = 2 * (a.warmup + a.steps) + 10
    weights = training = None
    t_setup = time.perf_counter()
    if a.ensemble:
        decks = lake.ensemble(a.ensemble, seed=7, side=(128, a.side_max), nts=nts)
        deck = decks[0]
        n_wet = sum(lake.n_wet_cells(d) for d in decks)
        g = GpuEnsemble(decks, precision=a.precision, preconditioner=a.precond, tol=a.tol)
        a.no_learned = True
        a.no_cpu_baseline = True
        workload = (f"ensemble of {a.ensemble} elliptic lakes, square sides 128-{a.side_max}, 20-80 layers, "
                    f"dx=100 m, dt=60 s, wind 0.05-0.3 Pa any direction, {a.precision}")
    elif a.river:
        deck = lake.channel(nx=a.nx, ny=a.ny, dx=100.0, depth=float(a.layers), shallow=a.layers / 3.0, dz=1.0,
                            dt=60.0, nts=nts, iturb=1, isal=1, temperature=20.0, ntr=a.tracers)
        tr = [1.0] * a.tracers
        deck = lake.open_boundary(deck, "W", 2, 2, a.ny + 1, lake.steady_series(a.river, 12.0, tr))
        deck = lake.open_boundary(deck, "E", 2, 2, a.ny + 1, lake.steady_series(a.river, 12.0, tr))
        n_wet = lake.n_wet_cells(deck)
        g = GpuLake(deck, precision=a.precision, preconditioner=a.precond, tol=a.tol)
        a.no_learned = True
        workload = (f"channel {a.nx}x{a.ny}x{a.layers}, dx=100 m, dt=60 s, {a.river:g} m3/s through open "
                    f"boundaries at both ends (12 C into 20 C), TVD + turbulence, plain PCG ({a.precond}) "
                    f"tol {a.tol:g}, {a.precision}")
        if a.tracers:
            workload += f", {a.tracers} tracers"
    else:
        deck = lake.ellipse(nx=a.nx, ny=a.ny, dx=a.dx, nlayers=a.layers, dt=a.dt, nts=nts, seed=1)
        if a.surfbc != 20:
            deck = lake.heat_budget(deck, a.surfbc)
        if a.tracers:
            km1 = len(deck["init_profile"])
            deck["tracer_profiles"] = np.stack([np.where(np.arange(km1) < km1 // (2 + q), 10.0, 2.0)
                                                for q in range(a.tracers)])
        n_wet = lake.n_wet_cells(deck)
        if a.precision != "fp64":
            a.no_learned = True  # the learned legs are measured against the plain FP64 leg
        if not a.no_learned:
            weights, training = train_on_random_winds(a, deck)
        t_setup = time.perf_counter()
        g = GpuLake(deck, precision=a.precision, preconditioner=a.precond, tol=a.tol)
        workload = (f"ellipse lake {a.nx}x{a.ny}x{a.layers}, dx={a.dx:g} m, dt={a.dt:g} s, TVD + "
                    f"turbulence, plain PCG ({a.precond}) tol {a.tol:g}, {a.precision}")
        if a.surfbc != 20:
            workload += f", surface heat budget ifsurfbc={a.surfbc}"
        if a.tracers:
            workload += f", {a.tracers} tracers"
    t_setup = time.perf_counter() - t_setup
    for n in range(1, a.warmup + 1):
        g.step(n)
    g.synchronize()
    st0 = g.stats()
    if not a.no_timers:
        g.timers(True)
    t0 = time.perf_counter()
    n0 = a.warmup + 1
    for n in range(n0, n0 + a.steps):
        g.step(n)
    g.synchronize()
    elapsed = time.perf_counter() - t0
    st1 = g.stats()
    ktimes = {} if a.no_timers else g.kernel_times()
    g.timers(False)

    value = n_wet * a.steps / elapsed
    iters = (st1["pcg_iters_total"] - st0["pcg_iters_total"]) / a.steps
    s_plain = g.get("s")[:g.lm]  # after warmup + steps, before the per-solve probe
    n_probe = 3
    trap_every_step = not a.ensemble and deck.get("itrap", 1) == 1
    per_solve = per_solve_iterations(g, deck, n0 + a.steps, n_probe) if trap_every_step else None
    elem = 8 if a.precision == "fp64" else 4

    roofline = None
    kernel_summary = {}
    if ktimes:
        best = None
        for name, (avg_ms, cnt) in ktimes.items():
            kernel_summary[name] = {"avg_ms": round(avg_ms, 5), "launches": cnt}
            if name not in ARRAYS_PER_CELL:
                continue
            tot = avg_ms * cnt
            if best is None or tot > best[1]:
                best = (name, tot, avg_ms)
        if best:
            name, _, avg_ms = best
            n_r, n_f64 = ARRAYS_PER_CELL[name]
            bytes_launch = (n_r * elem + n_f64 * 8) * n_wet
            achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
            traffic = None
            if not a.ensemble and not a.river and not a.tracers and a.surfbc == 20 and a.precision == "fp64":
                table = {(2048, 2048, 60, 50.0, 30.0): TRAFFIC_FP64_2048x2048x60_DX50}.get(
                    (a.nx, a.ny, a.layers, a.dx, a.dt), {})
                if name in table:
                    traffic = int(table[name] * n_wet)
            roofline = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic, "bytes_per_launch": bytes_launch,
                        "avg_launch_ms": round(avg_ms, 5),
                        "copy_GBs": None}
            cb = copy_bandwidth()
            roofline["copy_GBs"] = round(cb, 1) if cb else None
            roofline["stream_GBs"] = stream_bandwidth()

    tiles = g.tiles()
    lm, km1 = g.lm, g.km1
    learned = None
    if not a.no_learned:
        # the measured legs run one context at a time (each 2048 x 2048 FP64
        # context holds about half of the HBM)
        g.close()
        learned = {"training": training,
                   "eval_forcing": "the lake's own wind series (0.2 Pa sin(2 pi t / 24 h) from 30 deg), not in "
                                   "the training set",
                   "legs": [run_learned(a, deck, n_wet, weights, "fp64", s_plain, value),
                            run_learned(a, deck, n_wet, weights, "mixed", s_plain, value)]}
        if a.flex_period:
            learned["legs"].append(run_learned(a, deck, n_wet, weights, "fp64", s_plain, value, a.flex_period))

    cpu = None
    if not a.no_cpu_baseline:
        cdeck, cn_wet, what = deck, n_wet, "same lake"
        if a.nx * a.ny > 1024 * 512 and not (a.ensemble or a.river):
            cdeck = lake.ellipse(nx=a.cpu_side, ny=a.cpu_side, dx=a.dx, nlayers=a.layers, dt=a.dt,
                                 nts=a.cpu_steps + 3, seed=1)
            cn_wet = lake.n_wet_cells(cdeck)
            what = (f"the same generator on a {a.cpu_side}x{a.cpu_side} lake at dx={a.dx:g} m ({cn_wet} wet cells; "
                    f"the full lake takes minutes per step on one core)")
        if a.cpu_procs > 1:
            dt = cpu_baseline_procs(cdeck, a.cpu_steps, a.cpu_procs)
            cpu = {"value": round(a.cpu_procs * cn_wet * a.cpu_steps / dt, 1), "unit": "wet-cell-steps/s",
                   "cores": a.cpu_procs, "kind": "port",
                   "sample": f"FP64 oracle (C++ restatement), {a.cpu_procs} single-threaded processes, one copy of "
                             f"{what} each, steps 2-{a.cpu_steps + 1} (three fd calls each; step 1 untimed) timed "
                             f"together: {a.cpu_procs} x {a.cpu_steps} lake-steps in {dt:.1f} s (the slowest copy)"}
        else:
            dt, o = cpu_baseline(cdeck, a.cpu_steps)
            cpu = {"value": round(cn_wet * a.cpu_steps / dt, 1), "unit": "wet-cell-steps/s", "cores": 1,
                   "kind": "port",
                   "sample": f"FP64 oracle (C++ restatement, 1 thread), {what}, steps 2-{a.cpu_steps + 1} "
                             f"(three fd calls each; step 1 untimed), {dt:.1f} s"}

    out = {
        "metric": "wet-cell-steps/sec and PCG iters/step at 1 MI355X, learned vs plain, same error",
        "value": round(value, 1),
        "unit": "wet-cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "n_wet": n_wet, "lm": lm, "km1": km1, "parallelism": "1 GPU"},
        "pcg_iters_per_step": round(iters, 2),
        "pcg_iters_per_solve": per_solve,
        "pcg_relres_last": st1["pcg_relres_last"],
        "setup_s": round(t_setup, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned": learned,
        "tiles": tiles,
        "kernels": kernel_summary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
