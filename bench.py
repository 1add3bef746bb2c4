"""Throughput of the two-phase time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--repeats R]
                    [--workload large|swarm|rising] [--history-steps H]
                    [--no-cpu-baseline] [--cpu-steps N] [--no-learned]
                    [--no-ensemble] [--no-profile]

Workload (default, `large`): the bubble swarm on a 4096 x 8192 grid, 1 x 2
domain (walls in x along gravity, y periodic), 1024 bubbles of diameter
24-48 cells placed by seeded dart throwing with 4-cell gaps, Re=100, We=5,
Fr=1, density / viscosity ratios 1e-3 / 1e-2, fluid at rest.  A step is one
full time step (decide_dt + 3 Runge-Kutta stages).

The headline `value` is the chosen form of the step, mixed precision (FP32
interface and momentum stencils, FP64 vof / velocity updates, Poisson solve,
projection and every field in memory) with the learned curvature (its
network trained on the device first), in Mcell-steps/s = interior cells x
steps / wall seconds of the timed loop, with the state resident in HBM: the
median of R timed loops of K steps.  Beside it, from the same initial state:
the plain FP64 step (`fp64`), mixed precision without the network (`mixed`),
the speed-up over FP64, and the errors the chosen form is held to: the
rise-velocity history error max_t |u_r - u_r,FP64| / max_t |u_r,FP64| over
the first H steps, the relative gas-volume drift and the divergence.

`roofline` is for the kernel with the largest share of the headline step's
GPU time, timed with HIP events on the library stream around every launch of
a separate pass after the timed loops (the timed loops run without per-launch
events, by graph replay, as a production run does); its algorithmic bytes
per cell are BYTES_PER_CELL.  `cpu_baseline` times the CPU oracle (the FP64
C restatement of the reference step, OpenMP) built -O3 -march=native on this
machine, on a bounded sample of the same workload.  `ensemble` steps 256
independent rising bubbles of uneven grids as one batched ensemble.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from simplebubblyflowsolver_amd import configs, solver  # noqa: E402

METRIC = "Mcell-steps/s, mixed+learned vs plain FP64 at <=1% rise-velocity error"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per interior cell of one launch: every FP64 field
# the kernel must read or write, counted once.  The stage kernel of
# Runge-Kutta stage 0 (k_march*) reads vof, ux, uy, p, psi0 and writes vof',
# ux*, uy*, rhs and the three alpha sources (12 fields); with the previous
# stage's projection folded in (k_march*_corr) it also reads the previous
# potential and writes the pressure (stage 1: 9 reads + 8 writes = 17;
# stage 2: 9 + 7 = 16; mean 16.5).  The mixed-precision kernels store the
# same FP64 fields.  The staged kernels follow the reference's stages.
BYTES_PER_CELL = {
    "k_density": 16, "k_visux": 16, "k_visuy": 16, "k_dvof": 24, "k_normal_curv": 56,
    "k_force": 40, "k_fluxx": 48, "k_fluxy": 48, "k_vof_advance": 48, "k_mass_flux": 48,
    "k_stress": 56, "k_rhs_ux": 80, "k_rhs_uy": 72, "k_predict_ux": 56, "k_predict_uy": 56,
    "k_dct2_rhs": 40, "k_tdm_sequential": 16, "k_part_reduce": 16, "k_part_conv": 0,
    "k_dct3_psi": 16, "k_correct_ux": 40, "k_correct_uy": 40,
    "k_update_p": 24, "k_dt_partial": 16, "k_dt_final": 0,
    # fused path
    # (stages 1 and 2: 9 reads + 7 writes each; the mixed kernels keep the
    # momentum sources as FP32: 8 bytes fewer per cell in stage 0, and 16 and
    # 8 fewer in stages 1 and 2)
    "k_march": 12 * 8, "k_march_mixed": 12 * 8 - 8, "k_dct2_rows": 16,
    "k_post": 72, "k_march_corr": 16 * 8, "k_march_mixed_corr": 16 * 8 - 12,
    # learned curvature: the stage-0 gate reads the vof field
    "k_curv_gate": 8,
}

# HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE and
# WRITE_SIZE, scaled by a copy kernel of known size), read from the
# profile summaries; a line "traffic_bytes_per_launch: <kernel> <workload> <N>".
PROFILES = ROOT / "profiles"

# the learned curvature's training run (2^21 stencils)
TRAIN = dict(seed=2, nsamples=1 << 21, epochs=40, batch=4096, lr=1e-3)
MLP_FLOP_PER_SAMPLE = 3 * 2 * (26 * 64 + 64 * 64 + 64 * 1)  # forward + backward, 26-64-64-1
BF16_PEAK_TFLOPS = 2500.0


def measured_traffic(kernel: str, workload: str):
    for f in sorted(PROFILES.glob("*.md")):
        for line in f.read_text().splitlines():
            t = line.strip().strip("`").split()
            if len(t) == 4 and t[0] == "traffic_bytes_per_launch:" and t[1] == kernel and t[2] == workload:
                return float(t[3]), f.name
    return None, None


def workload(name: str) -> configs.Workload:
    if name == "large":
        return configs.large_swarm()
    if name == "swarm":
        return configs.bubble_swarm(2048, 4096, 64, steps=200, seed=1)
    if name == "rising":
        return configs.rising_bubble(128, 256, steps=2000)
    raise SystemExit(f"unknown workload {name}")


def profile_entries(lib) -> list:
    lib.sbfs_profile_entry.argtypes = [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                       C.POINTER(C.c_double)]
    out = []
    for k in range(lib.sbfs_profile_count()):
        name = C.create_string_buffer(128)
        cnt = C.c_int(0)
        ms = C.c_double(0.0)
        lib.sbfs_profile_entry(k, name, 128, C.byref(cnt), C.byref(ms))
        out.append((name.value.decode(), cnt.value, ms.value))
    return out


def kernel_table(prof, cells) -> list:
    out = []
    for n, cnt, ms in prof:
        bpc = BYTES_PER_CELL.get(n)
        avg_s = ms / cnt / 1e3
        out.append({"name": n, "launches": cnt, "total_ms": round(ms, 3), "avg_us": round(avg_s * 1e6, 2),
                    "alg_gbs": round(bpc * cells / avg_s / 1e9, 1) if bpc else None})
    return out


def run_leg(be, w, mixed: bool, learned: bool, a) -> dict:
    """One variant of the step from the workload's initial state: H untimed
    steps with the diagnostics after each (rise velocity, volume,
    divergence), W warm-up steps, R timed loops of K steps, and a separate
    per-kernel pass with HIP events."""
    lib = be.lib
    be.check(lib.sbfs_set_precision(1 if mixed else 0), "set_precision")
    net = solver.CurvatureNet(be) if learned else None
    if net:
        net.use(True)
    try:
        g = solver.Case(be, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
        try:
            g.load_initial_condition(w.vof())
            d0 = g.diagnostics()
            hist = []
            for _ in range(a.history_steps):
                g.integrate()
                hist.append(g.diagnostics())
            for _ in range(a.warmup):
                g.integrate()
            times = []
            for _ in range(a.repeats):
                be.check(lib.sbfs_synchronize(), "sync")
                t0 = time.perf_counter()
                for _ in range(a.steps):
                    g.integrate()
                be.check(lib.sbfs_synchronize(), "sync")
                times.append(time.perf_counter() - t0)
            d1 = g.diagnostics()
            cover = net.coverage(g) if net else None
            prof = []
            if not a.no_profile:
                be.check(lib.sbfs_profile_start(), "profile_start")
                for _ in range(a.profile_steps):
                    g.integrate()
                be.check(lib.sbfs_profile_stop(), "profile_stop")
                prof = profile_entries(lib)
        finally:
            g.close()
    finally:
        if net:
            net.use(False)
        be.check(lib.sbfs_set_precision(0), "set_precision")
    t = statistics.median(times)
    return {"t": t, "times": times, "d0": d0, "d1": d1, "hist": hist, "coverage": cover, "prof": prof,
            "value": w.cells * a.steps / t / 1e6}


def errors_vs(ref: dict, leg: dict) -> dict:
    ur_ref = np.array([d["rise_velocity"] for d in ref["hist"]])
    ur = np.array([d["rise_velocity"] for d in leg["hist"]])
    vols = np.array([d["volume"] for d in leg["hist"]] + [leg["d1"]["volume"]])
    v0 = leg["d0"]["volume"]
    return {"rise_velocity_history_rel_error": float(np.abs(ur - ur_ref).max() / max(np.abs(ur_ref).max(), 1e-300)),
            "history_steps": int(len(ur)),
            "volume_drift": float(np.abs(vols - v0).max() / v0),
            "divmax": max(d["divmax"] for d in leg["hist"] + [leg["d1"]])}


def train_network(be, nsamples: int, epochs: int) -> dict:
    """Train the learned curvature on the device; its gradient kernel is
    timed with HIP events on a short separate run for the MFMA rate."""
    lib = be.lib
    net = solver.CurvatureNet(be)
    be.check(lib.sbfs_profile_start(), "profile_start")
    net.train(seed=3, nsamples=min(nsamples, 1 << 18), epochs=1, batch=TRAIN["batch"], lr=1e-3)
    be.check(lib.sbfs_profile_stop(), "profile_stop")
    ent = {e[0]: e for e in profile_entries(lib)}
    t0 = time.perf_counter()
    hist = net.train(**dict(TRAIN, nsamples=nsamples, epochs=epochs))
    be.check(lib.sbfs_synchronize(), "sync")
    out = {"samples": nsamples, "epochs": epochs, "batch": TRAIN["batch"],
           "seconds": round(time.perf_counter() - t0, 3), "final_loss": float(hist[-1])}
    grad = ent.get("k_mlp_grad")
    if grad:
        avg = grad[2] / grad[1] / 1e3
        tf = MLP_FLOP_PER_SAMPLE * TRAIN["batch"] / avg / 1e12
        out["k_mlp_grad"] = {"avg_us": round(avg * 1e6, 2), "tflops": round(tf, 2),
                             "mfma_frac_of_bf16_peak": round(tf / BF16_PEAK_TFLOPS, 5)}
    return out


def ensemble_run(be, steps: int) -> dict:
    """The batched ensemble: 256 independent rising bubbles of uneven grids
    (configs.ensemble), mixed precision, band autotuned, `steps` timed steps
    after 6 untimed ones; Mcell-steps/s summed over the cases."""
    lib = be.lib
    ws = configs.ensemble(256, seed=3)
    cells = sum(w.nx * w.ny for w in ws)
    be.check(lib.sbfs_set_precision(1), "set_precision")
    try:
        e = solver.Ensemble(be, [dict(nx=w.nx, ny=w.ny, lx=w.lx, ly=w.ly, Re=w.Re, Fr=w.Fr, denr=w.denr,
                                      visr=w.visr, We=w.We) for w in ws])
        try:
            for c, w in zip(e.cases, ws):
                c.load_initial_condition(w.vof())
            for _ in range(2):
                e.step()
            band = e.autotune()
            for _ in range(4):
                e.step()
            be.check(lib.sbfs_synchronize(), "sync")
            t0 = time.perf_counter()
            for _ in range(steps):
                e.step()
            be.check(lib.sbfs_synchronize(), "sync")
            t = time.perf_counter() - t0
            vols = [c.diagnostics()["volume"] for c in e.cases[:8]]
        finally:
            e.close()
    finally:
        be.check(lib.sbfs_set_precision(0), "set_precision")
    sizes = {}
    for w in ws:
        sizes["%dx%d" % (w.nx, w.ny)] = sizes.get("%dx%d" % (w.nx, w.ny), 0) + 1
    return {"value": round(cells * steps / t / 1e6, 3), "unit": "Mcell-steps/s", "ms_per_step": round(t / steps * 1e3, 4),
            "steps": steps, "cases": len(ws), "cells": cells, "grids": sizes, "band_autotuned": band,
            "precision": "mixed", "finite": bool(np.all(np.isfinite(vols)))}


def host_threads() -> int:
    """The CPU cores this process may use: the affinity set (what `nproc`
    prints), capped by OMP_NUM_THREADS when the machine sets it (the share
    of the host that belongs to this GPU)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else n


def cpu_baseline(w: configs.Workload, steps: int) -> dict:
    """The oracle built -O3 -march=native here, OpenMP on host_threads(),
    1 warm-up step then `steps` timed steps of the same workload."""
    import oracle
    threads = host_threads()
    with tempfile.TemporaryDirectory() as tmp:
        lib = oracle.compile_library(Path(tmp) / "liboracle_native.so", opt=("-O3", "-march=native"))
        o_mod = oracle.Build(lib)
        o_mod.set_threads(threads)
        o_mod.set_poisson(dct=o_mod.DCT_FFT, tdm=o_mod.TDM_SEQUENTIAL)
        o = o_mod.new_case(w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
        try:
            o.load_initial_condition(w.vof())
            o.integrate()  # warm-up
            t0 = time.perf_counter()
            for _ in range(steps):
                o.integrate()
            t = time.perf_counter() - t0
        finally:
            o.close()
    return {"value": round(w.cells * steps / t / 1e6, 3), "unit": "Mcell-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} steps of {w.name} after 1 warm-up step: the CPU oracle (FP64 C restatement "
                      f"of the reference step) built gcc -O3 -march=native -ffp-contract=off, OpenMP on "
                      f"{threads} threads, {t:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=3, help="timed loops of --steps steps; the median is reported")
    ap.add_argument("--workload", default="large")
    ap.add_argument("--history-steps", type=int, default=25,
                    help="untimed steps from the initial state with diagnostics after each (error histories)")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="headline in mixed precision without the learned curvature (no training)")
    ap.add_argument("--train-samples", type=int, default=TRAIN["nsamples"],
                    help="training stencils of the learned curvature (fewer only for counter runs)")
    ap.add_argument("--train-epochs", type=int, default=TRAIN["epochs"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="buffer size of the copy-bandwidth measurement (0: skip)")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=10,
                    help="steps of the separate per-kernel timing pass")
    ap.add_argument("--no-ensemble", dest="ensemble", action="store_false")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the time step runs on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        return 2

    w = workload(a.workload)
    be = solver.product_backend()
    lib = be.lib
    be.set_dt_coefs(*configs.DT_COEFS)
    lib.sbfs_set_path(1)

    fp64 = run_leg(be, w, mixed=False, learned=False, a=a)
    print(f"fp64: {fp64['value']:.1f} Mcell-steps/s", file=sys.stderr, flush=True)
    mixed = run_leg(be, w, mixed=True, learned=False, a=a)
    print(f"mixed: {mixed['value']:.1f} Mcell-steps/s", file=sys.stderr, flush=True)
    train = None
    head = mixed
    if a.learned:
        train = train_network(be, a.train_samples, a.train_epochs)
        head = run_leg(be, w, mixed=True, learned=True, a=a)
        print(f"mixed+learned: {head['value']:.1f} Mcell-steps/s", file=sys.stderr, flush=True)

    copy_gbs = None
    if a.copy_mib > 0:
        g = C.c_double(0.0)
        lib.sbfs_copy_bandwidth.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        be.check(lib.sbfs_copy_bandwidth(a.copy_mib << 20, 20, C.byref(g)), "copy_bandwidth")
        copy_gbs = round(g.value, 1)

    roof = None
    if head["prof"]:
        prof = head["prof"]
        total = sum(e[2] for e in prof)
        name, cnt, ms = max(prof, key=lambda e: e[2])
        avg_s = ms / cnt / 1e3
        bpc = BYTES_PER_CELL.get(name)
        alg_bytes = bpc * w.cells
        achieved = alg_bytes / avg_s / 1e9
        traffic, src = measured_traffic(name, w.name)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": name, "avg_us": round(avg_s * 1e6, 2), "launches": cnt,
                "alg_bytes_per_launch": round(alg_bytes), "bytes_per_cell": round(bpc, 3),
                "share_of_gpu_time": round(ms / total, 3), "traffic_source": src,
                "measured_copy_gbs": copy_gbs,
                "frac_of_measured_copy": round(achieved / copy_gbs, 4) if copy_gbs else None}

    err = errors_vs(fp64, head)
    speedup = fp64["t"] / head["t"]
    out = {
        "metric": METRIC, "value": round(head["value"], 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(head["t"] / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": w.name, "grid": [w.nx, w.ny], "bubbles": len(w.bubbles), "Re": w.Re, "We": w.We,
                   "denr": w.denr, "visr": w.visr,
                   "precision": "mixed: FP32 interface and momentum stencils; FP64 vof / velocity updates, "
                                "Poisson solve, projection, reductions and every field in memory",
                   "curvature": "learned (BF16-MFMA MLP, plain fallback)" if a.learned else "plain",
                   "repeats": a.repeats, "timed_loops_s": [round(t, 5) for t in head["times"]]},
        "fp64": {"value": round(fp64["value"], 3), "ms_per_step": round(fp64["t"] / a.steps * 1e3, 4),
                 "timed_loops_s": [round(t, 5) for t in fp64["times"]],
                 "volume_drift": abs(fp64["d1"]["volume"] - fp64["d0"]["volume"]) / fp64["d0"]["volume"],
                 "divmax": fp64["d1"]["divmax"], "kernels": kernel_table(fp64["prof"], w.cells)},
        "speedup_vs_fp64": round(speedup, 3),
        "errors_vs_fp64": err,
        "bar": {"speedup_ge_2": speedup >= 2.0, "rise_velocity_le_1pct": err["rise_velocity_history_rel_error"] <= 0.01,
                "volume_drift_le_1e-10": err["volume_drift"] <= 1e-10},
        "cells_learned_fraction": head["coverage"],
        "mixed": {"value": round(mixed["value"], 3), "ms_per_step": round(mixed["t"] / a.steps * 1e3, 4),
                  "speedup_vs_fp64": round(fp64["t"] / mixed["t"], 3), "errors_vs_fp64": errors_vs(fp64, mixed)},
        "train": train,
        "roofline": roof,
        "kernels": kernel_table(head["prof"], w.cells),
    }
    if a.ensemble:
        out["ensemble"] = ensemble_run(be, a.ensemble_steps)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
