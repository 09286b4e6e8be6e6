#!/usr/bin/env python3
"""Benchmark of the fused D2Q9 MRT / smoothed-profile time step on one MI355X.

Workload: two disks settling in a Papanastasiou-Bingham fluid on a 401 x 1601
node grid (2d_nnf_lbm_amd.configs.two_disk_bingham).  A step is one pass of
the whole time step (streaming, bounce-back, smoothed-profile coupling, force
and torque, particle update, MRT/Papanastasiou collision) with the state
resident in HBM.  The headline value is the MLUPS of the FP32-collide /
FP16-stored path; the plain FP64 path is measured beside it, and the
reference program is timed on one host core as the CPU baseline.  Further
legs: the one-disk Newtonian FP64 case, the 504-disk BF16 cavity, the
256-cavity FP16 batch, the settling-velocity adjoint and the FP16-vs-FP64
accuracy run, each with the reference program of the same kind timed beside
it where one exists (the 504-disk program only with --cpu-dense).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--accuracy-steps A]
                       [--cpu-dense] [--no-cpu-baseline]
"""
import argparse
import importlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join("profiles", "k_step_two_disk_bingham.md")


def pmc_traffic(storage):
    """HBM bytes per launch of the fused step kernel from the committed
    rocprofv3 summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by what the
    counter read for a copy of known size); None when the summary is absent."""
    try:
        with open(os.path.join(REPO, TRAFFIC_SUMMARY)) as fh:
            text = fh.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch\s+(.*)", text)
    if not m:
        return None
    vals = dict(kv.split("=") for kv in m.group(1).split())
    return float(vals[storage]) if storage in vals else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measure(pkg, configs, storage, steps, warmup, workload=None):
    cfg, part = workload if workload is not None else configs.two_disk_bingham(storage)
    s = pkg.D2Q9(cfg)
    s.initialize()
    s.set_particles(part)
    s.step(0, warmup, record=False)
    s.sync()
    t0 = time.perf_counter()
    ms_total, ms_main = s.step_timed(warmup, steps)
    s.sync()
    wall = time.perf_counter() - t0
    n = configs.nodes(cfg)
    bpn = configs.bytes_per_node(storage, cfg["rheology"] == 1)
    per_launch = ms_main / steps  # ms per launch of the fused step kernel
    # the same steps with the per-step history (particle rows, mass and
    # yielded count: what the reference writes every step), by wall clock
    rec_steps = min(steps, 500)
    s.step(warmup + steps, rec_steps, record=True)  # sizes the history buffers
    s.sync()
    t0 = time.perf_counter()
    s.step(warmup + steps + rec_steps, rec_steps, record=True)
    s.sync()
    ms_rec = (time.perf_counter() - t0) * 1e3 / rec_steps
    out = dict(
        storage=storage,
        ms_per_step=ms_total / steps,
        ms_main_kernel=per_launch,
        wall_s=wall,
        mlups=n * steps / (ms_total * 1e-3) / 1e6,
        bytes_per_node=bpn,
        achieved_gbs=bpn * n / (per_launch * 1e-3) / 1e9,
        ms_per_step_recording=ms_rec,
        mlups_recording=n / (ms_rec * 1e-3) / 1e6,
    )
    s.close()
    return out


def measure_ensemble(pkg, configs, steps):
    """256 oscillating-cylinder cavities (sides U{128..1024}, Bn U[0,10],
    seed 1) in one batch: autotuned tiling, graph-replayed steps."""
    ens = configs.cavity_ensemble()
    nodes = sum(configs.nodes(c) for c, _ in ens)
    b = pkg.Batch([c for c, _ in ens])
    b.initialize()
    for i, (_, part) in enumerate(ens):
        b.set_particles(i, part)
    t0 = time.perf_counter()
    tiles = b.tune(steps=4)
    tune_s = time.perf_counter() - t0
    b.initialize()
    for i, (_, part) in enumerate(ens):
        b.set_particles(i, part)
    chunk = 16
    b.set_graph(chunk)
    b.step(0, chunk + 2, record=False)
    ms = b.step_timed(chunk + 2, steps)
    b.close()
    vals, counts = np.unique(tiles, return_counts=True)
    return dict(workload="256 lid- and side-wall-driven cavities, each with one cylinder "
                         "(D=side/8) in prescribed x-oscillation, sides U{128..1024}, "
                         "Bn U[0,10], FP32 collide / FP16 storage, one launch per step",
                simulations=len(ens), nodes_per_step=nodes, steps=steps, graph_chunk=chunk,
                mlups=nodes * steps / (ms * 1e-3) / 1e6, ms_per_step=ms / steps,
                achieved_gbs=configs.bytes_per_node("fp16", True) * nodes * steps / (ms * 1e-3) / 1e9,
                tiles_per_block={int(v): int(c) for v, c in zip(vals, counts)},
                tune_seconds=tune_s)


def measure_adjoint(pkg, configs, steps):
    """Gradient of the mean settling velocity over the last 512 steps with
    respect to (tau_y, mu0, rho_p): one disk, 513x1025 nodes, FP32 storage,
    checkpoints every 64 steps."""
    cfg, part = configs.settling_adjoint("fp32")
    s = pkg.D2Q9(cfg)
    s.initialize()
    s.set_particles(part)
    window, chk = min(512, steps), 64
    t0 = time.perf_counter()
    loss, grad, mem = s.adjoint_velocity(steps, window, chk)
    wall = time.perf_counter() - t0
    s.close()
    nodes = configs.nodes(cfg)
    return dict(workload="one disk (D=32, density ratio 1.2) settling in a 513x1025-node "
                         "Papanastasiou-Bingham channel; loss = mean V_y over the last %d steps"
                         % window,
                steps=steps, checkpoint_every=chk, seconds=wall,
                forward_equivalent_mlups=nodes * steps / wall / 1e6,
                loss=loss, grad_tau_y=grad[0], grad_mu0=grad[1], grad_rho_p=grad[2],
                device_bytes=mem)


def accuracy(pkg, configs, steps):
    """FP16-stored vs FP64 on the same run: terminal velocity, drag
    coefficient and yielded-area fraction after `steps` steps."""
    res = {}
    for storage in ("fp64", "fp16"):
        cfg, part = configs.two_disk_bingham(storage)
        s = pkg.D2Q9(cfg)
        s.initialize()
        s.set_particles(part)
        hist = s.step(0, steps)
        vy0, cd0 = configs.settling_metrics(hist, 2, cfg, 0)
        vy1, cd1 = configs.settling_metrics(hist, 2, cfg, 1)
        cdm = [configs.mean_drag_coefficient(hist, cfg, p) for p in (0, 1)]
        yielded = hist[-1][-1] / configs.nodes(cfg)
        res[storage] = dict(vy=[vy0, vy1], cd=[cd0, cd1], cdm=cdm, yielded=yielded)
        s.close()
    a, b = res["fp64"], res["fp16"]
    rel = lambda x, y: abs(x - y) / max(abs(y), 1e-300)
    return dict(
        steps=steps,
        terminal_velocity_rel_err=max(rel(b["vy"][i], a["vy"][i]) for i in range(2)),
        drag_coefficient_rel_err=max(rel(b["cd"][i], a["cd"][i]) for i in range(2)),
        drag_coefficient_mean_last_1000_rel_err=max(rel(b["cdm"][i], a["cdm"][i])
                                                    for i in range(2)),
        yielded_fraction_rel_err=rel(b["yielded"], a["yielded"]),
        fp64_vy=a["vy"], fp16_vy=b["vy"], fp64_yielded_fraction=a["yielded"],
    )


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The reference program (oracle/_ref/f6_bing_400x1600: the settling-disk
    Bingham program on the same 401 x 1601 node grid, 40 steps) timed on one
    host core; falls back to the C restatement of the oracle when that build is
    absent."""
    exe = os.path.join(REPO, "oracle", "_ref", "f6_bing_400x1600")
    nodes = 401 * 1601
    if os.path.exists(exe):
        steps = 40
        with tempfile.TemporaryDirectory(prefix="refrun_") as d:
            t0 = time.perf_counter()
            r = subprocess.run([exe], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                               text=True, timeout=600)
            wall = time.perf_counter() - t0
        m = re.search(r"RUNTIME=\s*([0-9.Ee+-]+)", r.stdout)
        cpu_t = float(m.group(1)) if m else wall
        out = dict(value=nodes * steps / wall / 1e6, unit="MLUPS", cores=1, kind="reference",
                   sample="reference Bingham settling-disk program (F6, patched grid "
                          "NX=400 NY=1600, one disk, per-step history I/O to a temporary "
                          "directory), %d steps, wall %.2f s, CPU_TIME %.2f s" % (steps, wall, cpu_t))
        # the reference has no threads: independent copies, one per core of
        # this process's share of the host (at most 16), run side by side
        copies = max(1, min(16, len(os.sched_getaffinity(0))))
        dirs = [tempfile.mkdtemp(prefix="refrun_") for _ in range(copies)]
        try:
            t0 = time.perf_counter()
            procs = [subprocess.Popen([exe], cwd=d, stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL) for d in dirs]
            codes = [p.wait(timeout=900) for p in procs]
            wall_n = time.perf_counter() - t0
        finally:
            for d in dirs:
                shutil.rmtree(d, ignore_errors=True)
        if all(c == 0 for c in codes):
            out["copies"] = dict(copies=copies, value=copies * nodes * steps / wall_n / 1e6,
                                 unit="MLUPS", wall_s=round(wall_n, 2))
        out["cpu_model"] = cpu_model()
        return out
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    import numpy as np
    oracle.build()
    pkg = importlib.import_module("2d_nnf_lbm_amd")
    cfg, part = pkg.configs.two_disk_bingham("fp64")
    cfg = dict(cfg, rate_mode=0)
    st = oracle.State(oracle.make_params(cfg), 2)
    st.initialize()
    st.set_particles(part)
    steps = 20
    t0 = time.perf_counter()
    st.run(0, steps, record=False)
    wall = time.perf_counter() - t0
    return dict(value=nodes * steps / wall / 1e6, unit="MLUPS", cores=1, kind="port",
                sample="C restatement of the reference step on the two-disk workload, "
                       "%d steps, wall %.2f s" % (steps, wall))


def run_reference(name, files=None, timeout=900):
    """Runs the reference timing program oracle/_ref/<name> once in a fresh
    temporary directory (with `files`, name -> text, written there first).
    Returns (wall seconds, the program's RUNTIME, the per-iteration CPU clock
    rows (KK, seconds since the program's START) it wrote to fort.98), or None
    when the build is absent or the run fails."""
    exe = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory(prefix="refrun_") as d:
        for fname, text in (files or {}).items():
            with open(os.path.join(d, fname), "w") as fh:
                fh.write(text)
        t0 = time.perf_counter()
        r = subprocess.run([exe], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                           text=True, timeout=timeout)
        wall = time.perf_counter() - t0
        if r.returncode:
            log("bench: %s exited with %d" % (name, r.returncode))
            return None
        m = re.search(r"RUNTIME=\s*([0-9.Ee+-]+)", r.stdout)
        clock = []
        path = os.path.join(d, "fort.98")
        if os.path.exists(path):
            with open(path) as fh:
                clock = [(int(a), float(b)) for a, b in (ln.split() for ln in fh if ln.strip())]
    return wall, (float(m.group(1)) if m else None), clock


def per_step_seconds(clock):
    """Seconds per time-loop iteration from the clock rows, leaving out the
    first iteration and everything before it (set-up, first touches)."""
    (k0, t0), (k1, t1) = clock[0], clock[-1]
    return (t1 - t0) / (k1 - k0)


def cpu_baseline_newtonian():
    """The reference Newtonian single-disk program (F6 with the one-disk
    workload's 200 x 600 cell grid and MRT rate set, 600 steps) on one host
    core."""
    res = run_reference("f6_newt_200x600")
    if res is None:
        return None
    wall, runtime, _ = res
    steps, nodes = 600, 201 * 601
    return dict(value=nodes * steps / wall / 1e6, unit="MLUPS", cores=1, kind="reference",
                sample="reference free-fall program (F6, Newtonian, NX=200 NY=600, alternative "
                       "MRT rates), %d steps, wall %.2f s (set-up included), CPU_TIME %.2f s"
                       % (steps, wall, runtime or float("nan")))


def cpu_baseline_ensemble(configs):
    """The reference oscillating-cylinder program (F4, Papanastasiou fluid)
    on square cavities of sides 128, 512 and 1024, one host core each; the
    per-node cost of one step, interpolated linearly in the side, summed over
    the 256 cavities of the batch gives the CPU time of one batch step."""
    sides, spn = [], []
    rows = {}
    for name, side in (("f4_cav_128", 128), ("f4_cav_512", 512), ("f4_cav_1024", 1024)):
        res = run_reference(name)
        if res is None or len(res[2]) < 2:
            return None
        sec = per_step_seconds(res[2])
        n = (side + 1) ** 2
        sides.append(side)
        spn.append(sec / n)
        rows[str(side)] = dict(steps_timed=res[2][-1][0] - res[2][0][0],
                               ms_per_step=sec * 1e3, mlups=n / sec / 1e6,
                               wall_s=round(res[0], 2))
    ens = configs.cavity_ensemble()
    nodes = [configs.nodes(c) for c, _ in ens]
    step_s = sum(n * float(np.interp(c["nx"], sides, spn)) for n, (c, _) in zip(nodes, ens))
    return dict(value=sum(nodes) / step_s / 1e6, unit="MLUPS", cores=1, kind="reference",
                sample="reference oscillating-cylinder program (F4, Papanastasiou fluid, "
                       "output off) on cavities of side 128, 512 and 1024, per-step CPU clock "
                       "after the first step; the batch's 256 sides priced by linear "
                       "interpolation of the per-node cost (extrapolated, one core)",
                seconds_per_batch_step=step_s, measured=rows)


def cpu_baseline_dense(configs):
    """The reference multi-particle program (F3) with the dense workload's
    504 disks on the 2048 x 2048 cell cavity, three steps on one host core;
    per-step cost from its CPU clock after the first step.  About 2.5 minutes
    and 17 GB of host memory (per-particle full-grid solid fractions)."""
    pos = np.loadtxt(configs.DENSE_FILE)
    text = "".join("%r %r\n" % (float(x), float(y)) for x, y in pos)
    res = run_reference("f3_dense_2048", {"0_Particle_Initial_Position.txt": text}, timeout=1200)
    if res is None or len(res[2]) < 2:
        return None
    sec = per_step_seconds(res[2])
    n = 2049 * 2049
    return dict(value=n / sec / 1e6, unit="MLUPS", cores=1, kind="reference",
                sample="reference multi-particle program (F3, 504 disks from the dense "
                       "workload's positions, NX=NY=2048), %d steps, per-step CPU clock after "
                       "the first step, wall %.1f s" % (res[2][-1][0], res[0]),
                seconds_per_step=sec)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--accuracy-steps", type=int, default=50000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-dense", action="store_true",
                    help="also time the reference multi-particle program on the 504-disk "
                         "cavity (about 2.5 minutes and 17 GB of host memory)")
    ap.add_argument("--newtonian-steps", type=int, default=2000,
                    help="timed steps of the one-disk Newtonian FP64 run (0 skips it)")
    ap.add_argument("--free-steps", type=int, default=200,
                    help="timed steps of the particle-free 2049x2049 FP16 Bingham cavity, the "
                         "tile path alone at a large grid (0 skips it)")
    ap.add_argument("--dense-steps", type=int, default=200,
                    help="timed steps of the 504-disk BF16 cavity (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=400,
                    help="timed steps of the 256-cavity FP16 batch (0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=4096,
                    help="steps of the settling-velocity adjoint run (0 skips it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one device-to-device copy of 9 FP64 planes (known byte "
                         "count) so hardware traffic counters can be scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the fused step runs on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    pkg = importlib.import_module("2d_nnf_lbm_amd")
    configs = importlib.import_module("2d_nnf_lbm_amd.configs")
    if args.copy_calibration:
        cfg, _ = configs.two_disk_bingham("fp64")
        s = pkg.D2Q9(cfg)
        s.copy_post()
        s.sync()
        plane = (configs.nodes(cfg) + 255) // 256 * 256
        log("bench: copy calibration, %d bytes read and written" % (9 * plane * 8))
        s.close()
    log("bench: fp64 leg")
    r64 = measure(pkg, configs, "fp64", args.steps, args.warmup)
    log("bench: fp16 leg")
    r16 = measure(pkg, configs, "fp16", args.steps, args.warmup)
    newt = None
    if args.newtonian_steps > 0:
        log("bench: one-disk newtonian, fp64")
        wl = configs.one_disk_newtonian("fp64")
        rn = measure(pkg, configs, "fp64", args.newtonian_steps, 100, workload=wl)
        newt = dict(workload="one disk (D=25, density ratio 1.5) falling from rest in a "
                             "Newtonian fluid (tau=0.6), 201x601 nodes, bounce-back walls, "
                             "plain FP64",
                    nodes=configs.nodes(wl[0]), steps=args.newtonian_steps,
                    mlups=rn["mlups"], ms_per_step=rn["ms_per_step"],
                    kernel_us_per_launch=rn["ms_main_kernel"] * 1e3,
                    achieved_gbs=rn["achieved_gbs"],
                    roofline_frac=rn["achieved_gbs"] / HBM_PEAK_GBS,
                    algorithmic_bytes_per_node=rn["bytes_per_node"])
    dense = None
    if args.dense_steps > 0:
        log("bench: 504-disk cavity, bf16")
        wl = configs.dense_sedimentation(0.1, "bf16")
        rd = measure(pkg, configs, "bf16", args.dense_steps, 20, workload=wl)
        dense = dict(workload="504 disks (D=16, density ratio 1.1) settling in a "
                              "Papanastasiou-Bingham fluid (Bn=0.1), 2049x2049 nodes, "
                              "FP32 collide / BF16 deviation storage",
                     nodes=configs.nodes(wl[0]), steps=args.dense_steps,
                     mlups=rd["mlups"], ms_per_step=rd["ms_per_step"],
                     kernel_us_per_launch=rd["ms_main_kernel"] * 1e3,
                     achieved_gbs=rd["achieved_gbs"],
                     roofline_frac=rd["achieved_gbs"] / HBM_PEAK_GBS,
                     algorithmic_bytes_per_node=rd["bytes_per_node"])
    free = None
    if args.free_steps > 0:
        log("bench: particle-free 2049x2049 cavity, fp16")
        cfg, part = configs.dense_sedimentation(0.1, "fp16")
        rf = measure(pkg, configs, "fp16", args.free_steps, 20, workload=(cfg, part[:, :0]))
        free = dict(workload="the 504-disk cavity's fluid (Papanastasiou-Bingham, Bn=0.1 "
                             "parameters) with no particles, at rest (every node takes the "
                             "small-rate branch), 2049x2049 nodes, FP32 collide / FP16 "
                             "deviation storage",
                    nodes=configs.nodes(cfg), steps=args.free_steps,
                    mlups=rf["mlups"], ms_per_step=rf["ms_per_step"],
                    kernel_us_per_launch=rf["ms_main_kernel"] * 1e3,
                    achieved_gbs=rf["achieved_gbs"],
                    roofline_frac=rf["achieved_gbs"] / HBM_PEAK_GBS,
                    algorithmic_bytes_per_node=rf["bytes_per_node"])
    ensemble = None
    if args.ensemble_steps > 0:
        log("bench: 256-cavity batch, fp16")
        ensemble = measure_ensemble(pkg, configs, args.ensemble_steps)
    adjoint = None
    if args.adjoint_steps > 0:
        log("bench: settling-velocity adjoint, fp32")
        adjoint = measure_adjoint(pkg, configs, args.adjoint_steps)
    log("bench: accuracy leg")
    acc = accuracy(pkg, configs, args.accuracy_steps) if args.accuracy_steps > 0 else None
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        cpu = cpu_baseline()
        if newt is not None:
            log("bench: cpu baseline, one-disk newtonian")
            newt["cpu_baseline"] = cpu_baseline_newtonian()
        if ensemble is not None:
            log("bench: cpu baseline, cavity programs")
            ensemble["cpu_baseline"] = cpu_baseline_ensemble(configs)
        if dense is not None and args.cpu_dense:
            log("bench: cpu baseline, 504-disk program")
            dense["cpu_baseline"] = cpu_baseline_dense(configs)
    cfg, _ = configs.two_disk_bingham("fp16")
    n = configs.nodes(cfg)
    line = {
        "metric": "MLUPS + % HBM roofline; plain FP64 vs FP16-stored at <=1e-3 rel err",
        "value": r16["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r16["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "two disks (D=40, density ratio 1.01) settling in a "
                        "Papanastasiou-Bingham fluid (Bn=0.5, tau=0.8, m=1e5), "
                        "401x1601 nodes, bounce-back walls, FP32 collide / FP16 deviation storage",
            "nodes": n,
            "storage": "fp16",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": r16["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": r16["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": pmc_traffic("fp16"),
            "traffic_source": TRAFFIC_SUMMARY,
            "algorithmic_bytes_per_launch": r16["bytes_per_node"] * n,
            "algorithmic_bytes_per_node": r16["bytes_per_node"],
            "kernel_us_per_launch": r16["ms_main_kernel"] * 1e3,
        },
        "cpu_baseline": cpu,
        "fp64": {
            "mlups": r64["mlups"],
            "ms_per_step": r64["ms_per_step"],
            "achieved_gbs": r64["achieved_gbs"],
            "roofline_frac": r64["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": pmc_traffic("fp64"),
            "algorithmic_bytes_per_node": r64["bytes_per_node"],
            "kernel_us_per_launch": r64["ms_main_kernel"] * 1e3,
            "recording": {"ms_per_step": r64["ms_per_step_recording"],
                          "mlups": r64["mlups_recording"]},
        },
        "fp16_recording": {"ms_per_step": r16["ms_per_step_recording"],
                           "mlups": r16["mlups_recording"],
                           "note": "the same FP16 steps with the per-step history "
                                   "(particle rows, mass, yielded count), wall clock"},
        "fp16_over_fp64": r16["mlups"] / r64["mlups"],
        "one_disk_newtonian_fp64": newt,
        "dense_sedimentation_bf16": dense,
        "particle_free_cavity_fp16": free,
        "cavity_ensemble_fp16": ensemble,
        "adjoint_fp32": adjoint,
        "accuracy_fp16_vs_fp64": acc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
