"""Benchmark: frames/s of the global FE-DIC hot path on one MI355X.

Workload (the 4 Mpx plain-GPU configuration): 2048x2048 synthetic speckle
(spot generator seed 0), ROI = centred 1800x1800 square minus a centred disc of
radius 300 px, structured Q4 mesh with 16 px elements (16,641 nodes, 11,430
elements; roi_mask passed, 2,923,392 pixels).  Frame t displacement
u(x,y,t) = t [eps (x - x0), -0.3 eps (y - y0)] + 1.5 px sin(2 pi x / W) t / 20 on
u_y, eps = 1e-3, plus N(0,1) grey noise.  Each step solves one frame
(GlobalCGSolver: cubic, max_iters=200, tol=1e-3, reg_strength=0.1) warm-started
from the previous frame's solution, then computes nodal Green-Lagrange strain
(DICMeshBased.run).  Images are resident in HBM before the timed region.

Prints ONE JSON line (see the keys below); --gpus > 1 is not supported (one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(H=2048, W=2048, side=1800, hole=300.0, e=16, seed=0, max_iters=200, tol=1e-3, reg=0.1,
                interp="cubic")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(n_frames: int):
    from d2ic_amd import speckle

    w = WORKLOAD
    fns = [speckle.affine_sinusoid(t, w["W"], w["H"]) for t in range(1, n_frames + 1)]
    ref, frames = speckle.speckle_sequence_device(w["H"], w["W"], w["seed"], fns, noise_sigma=1.0)
    mask = speckle.plate_with_hole_mask(w["H"], w["W"], w["side"], w["hole"])
    return ref, frames, mask


def measured_traffic():
    """HBM bytes per k_cg_persistent launch on this workload, from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary (scaled by a copy kernel), or None."""
    f = ROOT / "profiles" / "k_cg_persistent_plate2048_traffic.md"
    if not f.exists():
        return None
    import re

    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(ref_np, frames_np, mesh, mask, n_frames: int):
    """The oracle (a CPU restatement of the reference algorithm, FP32) on the host cores."""
    from oracle import oracle as O

    w = WORKLOAD
    t0 = time.perf_counter()
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref_np.shape, roi_mask=mask)
    t_assets = time.perf_counter() - t0
    u = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    iters = []
    t0 = time.perf_counter()
    for k in range(n_frames):
        u, _, n = O.cg_solve(u, ref_np, frames_np[k], a, alpha_reg=w["reg"], max_iter=w["max_iters"], tol=w["tol"],
                             interp=w["interp"], dtype=np.float32)
        iters.append(n)
    dt = time.perf_counter() - t0
    return dict(value=n_frames / dt, unit="frames/s", cores=O.num_threads(), kind="port",
                sample=f"first {n_frames} frame(s) of the same sequence, FP32 oracle (C, OpenMP), "
                       f"iterations {iters}, {dt:.1f} s; asset build {t_assets:.1f} s not included")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-frames", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (scales rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path (--gpus must be 1).",
              file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calibration} MiB 5 times")
        return 0

    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, mask_to_mesh_assets

    torch.cuda.init()
    w = WORKLOAD
    K, Wm = int(args.steps), int(args.warmup)
    log(f"[bench] rendering {K} frames of {w['H']}x{w['W']} speckle on the GPU")
    ref, frames, mask = build_workload(max(K, 1))
    mesh, assets = mask_to_mesh_assets(mask, w["e"])
    cfg = MeshDICConfig(max_iters=w["max_iters"], tol=w["tol"], reg_strength=w["reg"], save_history=False)
    solver = GlobalCGSolver(interpolation=w["interp"])
    dic = DICMeshBased(mesh, solver, cfg)
    t0 = time.perf_counter()
    dic.prepare(ref, assets)
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    info = solver.plan.info()
    Np, Nn = info["n_pixels"], info["n_nodes"]
    B_pass = 48 * Np + 16 * Nn  # algorithmic bytes of one pixel-pass (reference asset layout, FP32)
    log(f"[bench] prepared in {t_prep:.2f} s: Np={Np} Nn={Nn} tiles={info['n_tiles']}")

    # warmup: untimed solves of frame 1 from zero
    for _ in range(Wm):
        dic.set_initial_guess(None)
        dic.run(frames[0])
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    kt, kms, passes, iters = [], [], [], []
    u_prev = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        dic.set_initial_guess(u_prev)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        res = dic.run(frames[k])
        ev1.record(stream)
        u_prev = res.u_nodal
        kt.append((ev0, ev1))
        passes.append(solver.last_pass_count)
        kms.append(solver.plan.kernel_ms())
        iters.append(res.diagnostics.info["n_iters"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    solve_ms = [a.elapsed_time(b) for a, b in kt]
    fps = K / elapsed
    tot_bytes = B_pass * float(sum(passes))
    # algorithmic bytes of the persistent CG kernel / its measured launch time
    achieved = tot_bytes / (sum(kms) * 1e-3) / 1e9
    traffic = measured_traffic()
    log(f"[bench] {K} frames in {elapsed:.3f} s -> {fps:.2f} frames/s; iters {iters}; passes {passes}")

    cpu = None
    if not args.no_cpu_baseline and args.cpu_frames > 0:
        log(f"[bench] CPU baseline: oracle on {args.cpu_frames} frame(s)")
        ref_np = ref.cpu().numpy()
        fr_np = [f.cpu().numpy() for f in frames[: args.cpu_frames]]
        cpu = cpu_baseline(ref_np, fr_np, mesh, mask, args.cpu_frames)

    line = {
        "metric": "frames/s at 4 Mpx Q4-16px; p50/p99 latency under load vs plain; HBM GB/s",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plate-with-hole 2048x2048 speckle, Q4-16 mesh (16641 nodes), warm-started frame sequence, "
                        "plain one-solve-per-launch path",
            "image": [w["H"], w["W"]], "element_size_px": w["e"], "n_pixels": Np, "n_nodes": Nn,
            "interpolation": w["interp"], "max_iters": w["max_iters"], "tol": w["tol"], "reg_strength": w["reg"],
            "iterations_per_frame": iters, "pixel_passes_per_frame": passes,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_cg_persistent (one launch per frame)",
            "algorithmic_bytes_per_launch": B_pass * float(np.mean(passes)),
            "bytes_per_pass": B_pass,
            "mean_kernel_ms": float(np.mean(kms)),
            "mean_solve_ms": float(np.mean(solve_ms)),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    # release device objects before interpreter teardown (profilers finalise then)
    solver.close()
    del dic, solver, frames, ref, res, u_prev
    torch.cuda.synchronize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
