"""Throughput of the two-phase time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--path fused|staged]
                    [--workload swarm|rising|large] [--no-cpu-baseline]
                    [--no-mixed] [--no-learned] [--no-ensemble] [--no-profile]

Workload (default): plain FP64 bubble swarm, 2048 x 4096 cells on a 1 x 2
domain (walls in x, periodic y), 64 bubbles of diameter 24-48 cells placed by
seeded dart throwing, Re=35, We=10, Fr=1, density / viscosity ratios 1e-3 /
1e-2, fluid at rest.  A step is one full time step (decide_dt + 3 Runge-Kutta
stages).  `value` is interior cells x steps / wall seconds of the timed loop,
in Mcell-steps/s, with the state resident in HBM.

`roofline` is for the kernel with the largest share of GPU time, timed with
HIP events on the library stream around every launch of a separate pass of
--profile-steps steps that follows the timed loop (the timed loop itself runs
without per-launch events, as a production run does); its algorithmic bytes
per cell (distinct fields read + written once, FP64) are listed in
BYTES_PER_CELL.  `cpu_baseline` times the CPU oracle (the FP64 C
restatement, OpenMP) on a bounded sample of the same workload.

Beside the headline: `mixed` reruns the workload in mixed precision, and
`mixed.learned_curvature` with the learned curvature after training its
network on the device; `ensemble` steps 256 independent rising bubbles of
uneven grids as one batched ensemble.  `--workload large` runs the
4096 x 8192 swarm with 1024 bubbles (Re 100, We 5) instead.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from simplebubblyflowsolver_amd import configs, solver  # noqa: E402

METRIC = "Mcell-steps/s, mixed+learned vs plain FP64 at <=1% rise-velocity error"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per interior cell of one launch: every FP64 field
# the kernel must read or write, counted once.  k_march differs per
# Runge-Kutta stage (stage 0: reads vof, ux, uy, p, psi0 and writes vof',
# ux*, uy*, rhs and the three alpha sources = 12 fields; stage 1: + the
# three beta sources = 15; stage 2: 8 reads, writes vof', ux*, uy*, rhs,
# den[1], txy = 14), so its entry is the mean over the three launches of a
# step.  With the previous stage's projection folded into stages 1 and 2
# (k_march_corr), k_march is stage 0 alone (12 fields) and k_march_corr
# also reads the previous potential and writes the pressure (stage 1: 9
# reads + 8 writes = 17; stage 2: 9 + 7 = 16; mean 16.5).  The staged
# kernels follow the reference's stage boundaries.
BYTES_PER_CELL = {
    "k_density": 16, "k_visux": 16, "k_visuy": 16, "k_dvof": 24, "k_normal_curv": 56,
    "k_force": 40, "k_fluxx": 48, "k_fluxy": 48, "k_vof_advance": 48, "k_mass_flux": 48,
    "k_stress": 56, "k_rhs_ux": 80, "k_rhs_uy": 72, "k_predict_ux": 56, "k_predict_uy": 56,
    "k_dct2_rhs": 40, "k_tdm_sequential": 16, "k_part_reduce": 16, "k_part_conv": 0,
    "k_dct3_psi": 16, "k_correct_ux": 40, "k_correct_uy": 40,
    "k_update_p": 24, "k_dt_partial": 16, "k_dt_final": 0,
    # fused path
    "k_march": (12 + 15 + 14) * 8 / 3, "k_march_mixed": (12 + 15 + 14) * 8 / 3, "k_pre": (10 + 13 + 12) * 8 / 3, "k_dct2_rows": 16,
    "k_post": 72, "k_march_corr": (17 + 16) * 8 / 2, "k_march_mixed_corr": (17 + 16) * 8 / 2,
}


def bytes_per_cell(name: str, names) -> float | None:
    """BYTES_PER_CELL, with k_march counted as stage 0 alone when the
    folded-projection kernels ran beside it."""
    if name in ("k_march", "k_march_mixed") and (name + "_corr") in names:
        return 12 * 8
    return BYTES_PER_CELL.get(name)

# HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE and
# WRITE_SIZE, scaled by a copy kernel of known size), read from the
# profile summaries; a line "traffic_bytes_per_launch: <kernel> <workload> <N>".
PROFILES = ROOT / "profiles"


def measured_traffic(kernel: str, workload: str):
    for f in sorted(PROFILES.glob("*.md")):
        for line in f.read_text().splitlines():
            t = line.strip().strip("`").split()
            if len(t) == 4 and t[0] == "traffic_bytes_per_launch:" and t[1] == kernel and t[2] == workload:
                return float(t[3]), f.name
    return None, None


def workload(name: str) -> configs.Workload:
    if name == "swarm":
        return configs.bubble_swarm(2048, 4096, 64, steps=200, seed=1)
    if name == "rising":
        return configs.rising_bubble(128, 256, steps=2000)
    if name == "large":
        return configs.large_swarm()
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


def cpu_baseline(w: configs.Workload, steps: int, threads: int) -> dict:
    import oracle
    oracle.build()
    oracle.set_threads(threads)
    oracle.set_poisson(dct=oracle.DCT_FFT, tdm=oracle.TDM_SEQUENTIAL)
    o = oracle.new_case(w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    o.load_initial_condition(w.vof())
    o.integrate()  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        o.integrate()
    t = time.perf_counter() - t0
    o.close()
    return {"value": round(w.cells * steps / t / 1e6, 3), "unit": "Mcell-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} steps of {w.name} after 1 warm-up step, "
                      f"CPU oracle (FP64 C restatement, OpenMP, {threads} threads), {t:.1f} s"}


def mixed_run(be, w, warmup, steps, t_fp64, d0, d1_fp64) -> dict:
    """The same workload from the same initial state with the mixed-precision
    stage kernel: throughput, speedup over the FP64 loop just timed, and the
    rise-velocity and gas-volume errors after the same number of steps."""
    lib = be.lib
    be.check(lib.sbfs_set_precision(1), "set_precision")
    try:
        g = solver.Case(be, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
        g.load_initial_condition(w.vof())
        for _ in range(warmup):
            g.integrate()
        be.check(lib.sbfs_synchronize(), "sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            g.integrate()
        be.check(lib.sbfs_synchronize(), "sync")
        t = time.perf_counter() - t0
        d = g.diagnostics()
        g.close()
    finally:
        be.check(lib.sbfs_set_precision(0), "set_precision")
    ur = d1_fp64["rise_velocity"]
    return {"value": round(w.cells * steps / t / 1e6, 3), "ms_per_step": round(t / steps * 1e3, 4),
            "speedup_vs_fp64": round(t_fp64 / t, 3),
            "rise_velocity_rel_error": abs(d["rise_velocity"] - ur) / max(abs(ur), 1e-300),
            "volume_drift": abs(d["volume"] - d0["volume"]) / d0["volume"], "divmax": d["divmax"],
            "precision": "FP32 interface and momentum stencils; FP64 vof / velocity updates, "
                         "Poisson solve, projection and all fields in memory",
            "learned_curvature": None}  # filled by the caller


# the learned curvature's training run (the chosen form's 2^21 stencils)
TRAIN = dict(seed=2, nsamples=1 << 21, epochs=40, batch=4096, lr=1e-3)
MLP_FLOP_PER_SAMPLE = 3 * 2 * (25 * 64 + 64 * 64 + 64 * 1)  # forward + backward, 25-64-64-1
BF16_PEAK_TFLOPS = 2500.0


def learned_run(be, w, warmup, steps, t_fp64, d0, d1_fp64) -> dict:
    """Mixed precision with the learned curvature. The network is trained on
    the device first, with its training kernel timed by HIP events for the
    MFMA rate; then the workload runs from the same initial state."""
    lib = be.lib
    net = solver.CurvatureNet(be)
    be.check(lib.sbfs_profile_start(), "profile_start")
    net.train(seed=3, nsamples=1 << 18, epochs=1, batch=TRAIN["batch"], lr=1e-3)
    be.check(lib.sbfs_profile_stop(), "profile_stop")
    grad = {e[0]: e for e in profile_entries(lib)}.get("k_mlp_grad")
    t0 = time.perf_counter()
    hist = net.train(**TRAIN)
    be.check(lib.sbfs_synchronize(), "sync")
    train_s = time.perf_counter() - t0
    be.check(lib.sbfs_set_precision(1), "set_precision")
    net.use(True)
    try:
        g = solver.Case(be, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
        g.load_initial_condition(w.vof())
        for _ in range(warmup):
            g.integrate()
        be.check(lib.sbfs_synchronize(), "sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            g.integrate()
        be.check(lib.sbfs_synchronize(), "sync")
        t = time.perf_counter() - t0
        d = g.diagnostics()
        cover = net.coverage(g)
        be.check(lib.sbfs_profile_start(), "profile_start")
        for _ in range(3):
            g.integrate()
        be.check(lib.sbfs_profile_stop(), "profile_stop")
        ent = {e[0]: e for e in profile_entries(lib)}
        curv, gate = ent.get("k_curv_mlp"), ent.get("k_curv_gate")
        g.close()
    finally:
        net.use(False)
        be.check(lib.sbfs_set_precision(0), "set_precision")
    ur = d1_fp64["rise_velocity"]
    out = {"value": round(w.cells * steps / t / 1e6, 3), "ms_per_step": round(t / steps * 1e3, 4),
           "speedup_vs_fp64": round(t_fp64 / t, 3),
           "rise_velocity_rel_error": abs(d["rise_velocity"] - ur) / max(abs(ur), 1e-300),
           "volume_drift": abs(d["volume"] - d0["volume"]) / d0["volume"], "divmax": d["divmax"],
           "cells_learned_fraction": cover,
           "train": {"samples": TRAIN["nsamples"], "epochs": TRAIN["epochs"], "batch": TRAIN["batch"],
                     "seconds": round(train_s, 3), "final_mse": float(hist[-1])}}
    if grad:
        avg = grad[2] / grad[1] / 1e3
        tf = MLP_FLOP_PER_SAMPLE * TRAIN["batch"] / avg / 1e12
        out["train"]["k_mlp_grad"] = {"avg_us": round(avg * 1e6, 2), "tflops": round(tf, 2),
                                      "mfma_frac_of_bf16_peak": round(tf / BF16_PEAK_TFLOPS, 5)}
    if curv:
        out["k_curv_mlp_avg_us"] = round(curv[2] / curv[1] * 1e3, 2)
    if gate:
        out["k_curv_gate_avg_us"] = round(gate[2] / gate[1] * 1e3, 2)
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--path", choices=["fused", "staged"], default="fused")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 interface and momentum stencils in the stage kernel")
    ap.add_argument("--workload", default="swarm")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="buffer size of the copy-bandwidth measurement (0: skip)")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false",
                    help="skip the mixed-precision comparison run")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip the learned-curvature run (training and stepping)")
    ap.add_argument("--no-ensemble", dest="ensemble", action="store_false",
                    help="skip the batched-ensemble run")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="steps of the separate per-kernel timing pass")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the time step runs on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        return 2

    w = workload(a.workload)
    be = solver.product_backend()
    lib = be.lib
    be.set_dt_coefs(*configs.DT_COEFS)
    lib.sbfs_set_path(1 if a.path == "fused" else 0)
    be.check(lib.sbfs_set_precision(1 if a.precision == "mixed" else 0), "set_precision")
    g = solver.Case(be, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    g.load_initial_condition(w.vof())
    d0 = g.diagnostics()
    for _ in range(a.warmup):
        g.integrate()
    be.check(lib.sbfs_synchronize(), "sync")
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g.integrate()
    be.check(lib.sbfs_synchronize(), "sync")
    t = time.perf_counter() - t0
    d1 = g.diagnostics()
    mixed = None
    if a.mixed and a.precision == "fp64" and a.path == "fused":
        mixed = mixed_run(be, w, a.warmup, a.steps, t, d0, d1)
        if a.learned:
            mixed["learned_curvature"] = learned_run(be, w, a.warmup, a.steps, t, d0, d1)
    # per-kernel times: a separate pass with HIP events around every launch
    # (the events keep the step out of graph replay, so it is not the timed loop)
    prof = []
    if not a.no_profile:
        be.check(lib.sbfs_profile_start(), "profile_start")
        for _ in range(min(a.steps, a.profile_steps)):
            g.integrate()
        be.check(lib.sbfs_profile_stop(), "profile_stop")
        prof = profile_entries(lib)
    g.close()

    value = w.cells * a.steps / t / 1e6
    copy_gbs = None
    if a.copy_mib > 0:
        g = C.c_double(0.0)
        lib.sbfs_copy_bandwidth.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        be.check(lib.sbfs_copy_bandwidth(a.copy_mib << 20, 20, C.byref(g)), "copy_bandwidth")
        copy_gbs = round(g.value, 1)
    roof = None
    kernels = []
    if prof:
        total = sum(e[2] for e in prof)
        names = {e[0] for e in prof}
        for n, cnt, ms in prof:
            bpc = bytes_per_cell(n, names)
            avg_s = ms / cnt / 1e3
            kernels.append({"name": n, "launches": cnt, "total_ms": round(ms, 3),
                            "avg_us": round(avg_s * 1e6, 2),
                            "alg_gbs": round(bpc * w.cells / avg_s / 1e9, 1) if bpc else None})
        name, cnt, ms = max(prof, key=lambda e: e[2])
        avg_s = ms / cnt / 1e3
        bpc = bytes_per_cell(name, names)
        alg_bytes = bpc * w.cells
        achieved = alg_bytes / avg_s / 1e9
        traffic, src = measured_traffic(name, w.name)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "kernel": name, "avg_us": round(avg_s * 1e6, 2), "launches": cnt,
                "alg_bytes_per_launch": round(alg_bytes), "bytes_per_cell": round(bpc, 3),
                "share_of_gpu_time": round(ms / total, 3),
                "traffic_source": src,
                "measured_copy_gbs": copy_gbs,
                "frac_of_measured_copy": round(achieved / copy_gbs, 4) if copy_gbs else None}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(t / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": w.name, "grid": [w.nx, w.ny], "bubbles": len(w.bubbles),
                   "Re": w.Re, "We": w.We,
                   "precision": "plain FP64" if a.precision == "fp64" else "mixed (FP32 stencils, FP64 rest)",
                   "path": a.path,
                   "volume_drift": abs(d1["volume"] - d0["volume"]) / d0["volume"],
                   "divmax": d1["divmax"]},
        "roofline": roof,
        "kernels": kernels,
        "mixed": mixed,
    }
    if a.ensemble and a.path == "fused":
        out["ensemble"] = ensemble_run(be, a.ensemble_steps)
    if not a.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_steps, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
