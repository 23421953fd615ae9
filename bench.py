"""Benchmark: the springy step of a 1024 x 1024 (1,048,576-vertex) grid cloth draping over a
sphere collider on one MI355X.

One "step" is one Simulation::step of the whole cloth (forces, RK4, collision, pins) at
dt = 1/600 s.  The cloth's state is resident in HBM before the timed region; the timed region is
exactly K steps bracketed by device synchronisations.

Prints ONE JSON line:
  value           simulated seconds per wall second (K * dt / wall time)
  roofline        the step kernel against HBM bandwidth: 48 algorithmic bytes per vertex-step
                  (x, y, z, vx, vy, vz read once and written once, f32) over the kernel's average
                  duration measured with HIP events on the simulation's stream
  cpu_baseline    the oracle (a single-threaded C++ restatement of the reference step) on a bounded
                  sample of the same workload, on this host's cores
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_VERTEX_STEP = 48
# HBM bytes per launch of the default step kernel (grid_tiled_kernel<ExactMath, 16, 16>) on the
# 1024 x 1024 cloth, from rocprofv3 FETCH_SIZE (13,249 KiB x 2.000) and WRITE_SIZE (24,576 KiB x
# 1.000), each scaled by a 256 MiB device copy's reading (profiles/grid_tiled_exact_cloth1024.md)
MEASURED_TRAFFIC = {(1024, "exact", "auto", "default"): (13249 * 2.0 + 24576 * 1.0) * 1024}
# VALU issue of the same kernel from rocprofv3 SQ_INSTS_VALU (per wave) and GRBM_GUI_ACTIVE (per XCD):
# 2,177 instructions x 4 cycles x 16 waves per SIMD against 140.8 k active cycles
# (profiles/grid_tiled_cloth1024_pmc.md); the step is bound by VALU issue, not by HBM
MEASURED_VALU = {(1024, "exact", "auto", "default"): {"insts_per_wave": 2177, "issue_cycles_per_simd": 139328,
                                                      "active_cycles": 140839, "issue_frac": 139328 / 140839}}
# the same for patch_kernel<ExactMath> on the 64-icosphere workload (FETCH_SIZE 216,682 KiB x 2.000,
# WRITE_SIZE 50,318 KiB x 1.000; profiles/batched_and_surfaces.md)
MEASURED_TRAFFIC_SURFACES = {"exact": (216682 * 2.0 + 50318 * 1.0) * 1024}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--n", type=int, default=1024, help="cloth side (vertices)")
    ap.add_argument("--math", choices=["exact", "fast"], default="exact")
    ap.add_argument("--kernel", choices=["auto", "grid", "grid_simple", "csr"], default="auto")
    ap.add_argument("--tile", choices=["default", "15x15", "7x31", "16x16"], default="default")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--workload", choices=["cloth", "batched", "surfaces", "rigid"], default="cloth",
                    help="cloth: one 1024x1024 cloth + sphere (default); batched: 4096 cloths of uneven sizes; "
                         "surfaces: 64 perturbed icospheres (2.1M vertices) on a ground plane; rigid: 1M unit "
                         "cubes tumbling onto a ground plane")
    ap.add_argument("--chunk", type=int, default=100, help="steps per captured HIP graph (batched)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "feriphys_amd", "weights", "cloth_surrogate.npy"),
                    help="surrogate network for the cloth workload's surrogate leg ('' to skip it)")
    ap.add_argument("--sur-shape", type=int, choices=[32, 64], default=32,
                    help="vertices per wave of the surrogate network kernel")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the library has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import feriphys_amd as fa
    from feriphys_amd import workloads as wl

    if args.workload == "batched":
        return bench_batched(args, np, fa, wl)
    if args.workload == "surfaces":
        return bench_surfaces(args, np, fa, wl)
    if args.workload == "rigid":
        return bench_rigid(args, np, fa, wl)
    scene = wl.config1(n=args.n, steps=args.steps)
    sim = wl.build_gpu(scene)
    kern = {"auto": fa.FP_KERNEL_AUTO, "grid": fa.FP_KERNEL_GRID, "grid_simple": fa.FP_KERNEL_GRID_SIMPLE,
            "csr": fa.FP_KERNEL_CSR}[args.kernel]
    sim.set_kernel(kern, fa.FP_MATH_EXACT if args.math == "exact" else fa.FP_MATH_FAST)
    if args.tile != "default":
        sim.set_tile_shape(0, {"15x15": 0, "7x31": 1, "16x16": 2}[args.tile])
    N = scene.cloths[0].rows * scene.cloths[0].cols
    dt = scene.config.dt

    # Warm up on the same scene, then restore its initial state so that the timed region is
    # exactly steps 1..K of the workload (the restore is outside the timed region).
    initial = sim.get_points(0)
    if args.warmup:
        sim.step(args.warmup)
        sim.set_points(0, initial)
    sim.synchronize()
    t0 = time.perf_counter()
    kernel_ms = sim.time_steps(args.steps)  # HIP events on the simulation's stream around the K steps
    sim.synchronize()
    wall = time.perf_counter() - t0

    pts = sim.get_points(0)
    if not np.isfinite(pts).all():
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)

    sim_s_per_s = args.steps * dt / wall
    ms_per_step = wall * 1e3 / args.steps
    kernel_avg_s = kernel_ms * 1e-3 / args.steps
    achieved = BYTES_PER_VERTEX_STEP * N / kernel_avg_s / 1e9

    # the same 600 steps with the fast math policy (one-step forces within 1e-5 of the exact ones;
    # 600-step trajectories checked against the f64 oracle in the parity tests)
    fast = None
    if args.math == "exact":
        sim.set_kernel(kern, fa.FP_MATH_FAST)
        sim.set_points(0, initial)
        sim.step(max(1, args.warmup))
        sim.set_points(0, initial)
        sim.synchronize()
        tf = time.perf_counter()
        fast_ms = sim.time_steps(args.steps)
        sim.synchronize()
        fwall = time.perf_counter() - tf
        xf = sim.get_points(0)[:, 1:4].astype(np.float64)
        favg = fast_ms * 1e-3 / args.steps
        fast = {"sim_s_per_s": args.steps * dt / fwall, "kernel_avg_us": favg * 1e6,
                "hbm_frac": BYTES_PER_VERTEX_STEP * N / favg / 1e9 / HBM_PEAK_GBS,
                "max_position_difference_vs_exact_m": float(np.abs(xf - pts[:, 1:4]).max())}
        sim.set_kernel(kern, fa.FP_MATH_EXACT)

    sur = surrogate_leg(args, np, fa, wl, kern, pts, initial, N, dt, sim_s_per_s)

    # PCIe-inclusive rate (never the headline): a caller that reads the 10-float records back after
    # every step, as the reference's demo renders every frame; host clock over 20 frames
    sim.set_points(0, initial)
    sim.synchronize()
    frames = 20
    tr = time.perf_counter()
    for _ in range(frames):
        sim.step(1)
        sim.get_points(0)
    frame_s = (time.perf_counter() - tr) / frames
    with_readback = {"ms_per_frame": frame_s * 1e3, "sim_s_per_s": dt / frame_s,
                     "readback_bytes_per_frame": 40 * N, "frames": frames}

    cpu = None
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        o = orc.build_scene(wl.config1(n=args.n, steps=args.cpu_steps))
        o.step(1)  # warm-up
        t = time.perf_counter()
        o.step(args.cpu_steps)
        cpu_wall = time.perf_counter() - t
        cpu = {
            "value": args.cpu_steps * dt / cpu_wall,
            "unit": "sim-s/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same {args.n}x{args.n} cloth + sphere after 1 warm-up step "
                      f"(single-threaded C++ restatement of the reference step, -O2, no FMA); host "
                      f"{platform.processor() or platform.machine()}, nproc {os.cpu_count()}",
            "vertex_steps_per_s": args.cpu_steps * N / cpu_wall,
        }

    # the headline is the fastest form that meets the error bar (the plain step has no error)
    use_sur = sur is not None and sur["rms_error_over_extent"] <= 1e-3 and sur["sim_s_per_s"] > sim_s_per_s
    line = {
        "metric": "sim-s/s, 1M-vert cloth: surrogate vs plain @<=1e-3 pos err; %HBM peak",
        "value": sur["sim_s_per_s"] if use_sur else sim_s_per_s,
        "value_source": "surrogate" if use_sur else "plain",
        "plain_sim_s_per_s": sim_s_per_s,
        "unit": "sim-s/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain springy step, {args.n}x{args.n} grid cloth ({N} vertices), top edge pinned, "
                               f"320-face icosphere collider, RK4 dt=1/600 s",
                   "vertices": N, "dt": dt, "kernel": args.kernel, "math": args.math},
        "vertex_steps_per_s": args.steps * N / wall,
        "fast_math": fast,
        "surrogate": sur,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((args.n, args.math, args.kernel, args.tile)),
                     "kernel_avg_us": kernel_avg_s * 1e6, "bytes_per_launch": BYTES_PER_VERTEX_STEP * N},
        "valu_issue": MEASURED_VALU.get((args.n, args.math, args.kernel, args.tile)),
        "with_readback": with_readback,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


MLP_FLOP_PER_VERTEX = 2 * (72 * 128 + 128 * 128 + 128 * 128 + 128 * 3)  # 84,736
MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_leg(args, np, fa, wl, kern, ref, initial, N, dt, plain_sim_s_per_s):
    """The same cloth stepped by the learned surrogate (8x steps, per-mesh fallback on) for the same
    simulated time, timed on the host clock around the whole run (it reads each big step's
    diagnostics back); the network kernel's launches timed with HIP events.  Error: the final
    positions against the plain run's."""
    if not args.weights or not os.path.exists(args.weights):
        return None
    params = np.load(args.weights, allow_pickle=False)
    net = fa.Surrogate(fa.SurrogateConfig(), params=params)
    ratio = net.config.ratio
    n_big = args.steps // ratio
    sim = wl.build_gpu(wl.config1(n=args.n, steps=args.steps))
    sim.set_kernel(kern, fa.FP_MATH_EXACT if args.math == "exact" else fa.FP_MATH_FAST)
    sim.set_surrogate(net)
    sim.set_surrogate_shape(args.sur_shape)
    if args.warmup:
        sim.step_surrogate(max(1, args.warmup // ratio))
        sim.set_points(0, initial)
    sim.synchronize()
    total_ms, mlp_ms, st = sim.time_surrogate(n_big)
    x = sim.get_points(0)[:, 1:4].astype(np.float64)
    xr = ref[:, 1:4].astype(np.float64)
    x0 = initial[:, 1:4].astype(np.float64)
    extent = (args.n - 1) * 0.05
    e = np.linalg.norm(x - xr, axis=1)
    disp = float(np.sqrt(np.mean(np.sum((xr - x0) ** 2, axis=1))))
    sim_s = n_big * ratio * dt / (total_ms * 1e-3)
    launches = n_big - st.plain_only  # big steps that ran the network
    mlp_avg_s = mlp_ms * 1e-3 / max(1, launches)
    tflops = MLP_FLOP_PER_VERTEX * N / mlp_avg_s / 1e12 if mlp_avg_s > 0 else None
    return {
        "sim_s_per_s": sim_s, "speedup_vs_plain": sim_s / plain_sim_s_per_s, "big_steps": n_big, "ratio": ratio,
        "ms_per_big_step": total_ms / max(1, n_big),
        "rms_error_over_extent": float(np.sqrt(np.mean(e * e))) / extent,
        "max_error_over_extent": float(e.max()) / extent,
        "rms_error_over_rms_displacement": float(np.sqrt(np.mean(e * e))) / disp if disp > 0 else None,
        "finite": bool(np.isfinite(x).all()),
        "fallbacks": st.fallbacks, "plain_only": st.plain_only, "network_launches": launches,
        "max_strain_residual": st.max_strain_residual,
        "max_energy_residual": st.max_energy_residual,
        "error_definition": "final positions vs the plain dt=1/600 run after the same simulated time; "
                            f"extent = cloth width {extent:.2f} m",
        "mlp_roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tflops / MFMA_BF16_PEAK_TFLOPS if tflops else None,
                         "kernel_avg_us": mlp_avg_s * 1e6, "flop_per_launch": MLP_FLOP_PER_VERTEX * N},
    }


def bench_batched(args, np, fa, wl):
    """4096 cloths (sides log-uniform in [8, 256], ~19.7M vertices), each with its own Config, one
    tiled launch per tile shape per step, steps replayed from HIP graphs of `chunk` steps."""
    scene = wl.config2(steps=args.steps)
    sim = wl.build_gpu(scene)
    sim.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT if args.math == "exact" else fa.FP_MATH_FAST)
    N = sum(c.rows * c.cols for c in scene.cloths)
    dt = scene.config.dt
    chosen = sim.autotune(10)
    initial = [sim.get_points(k) for k in range(len(scene.cloths))]
    if args.warmup:
        sim.step_graph(args.warmup, args.chunk)
        for k, rec in enumerate(initial):
            sim.set_points(k, rec)
    sim.synchronize()
    t0 = time.perf_counter()
    kernel_ms = sim.time_steps_graph(args.steps, args.chunk)
    sim.synchronize()
    wall = time.perf_counter() - t0
    kernel_avg_s = kernel_ms * 1e-3 / args.steps
    achieved = BYTES_PER_VERTEX_STEP * N / kernel_avg_s / 1e9

    cpu = None
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        sub = wl.Scene(scene.cloths[:64], scene.config, [], scene.jitter, scene.seed, args.cpu_steps)
        sub.mesh_winds = scene.mesh_winds[:64]
        sims = orc.build_scene_per_mesh(sub)
        n_sub = sum(c.rows * c.cols for c in sub.cloths)
        t = time.perf_counter()
        for o in sims:
            o.step(args.cpu_steps)
        cpu_wall = time.perf_counter() - t
        vsps = args.cpu_steps * n_sub / cpu_wall
        cpu = {"value": vsps / N * dt, "unit": "sim-s/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the first 64 cloths ({n_sub} vertices), scaled by vertex count "
                         f"to the {N}-vertex batch (single-threaded C++ restatement of the reference step, one "
                         f"Simulation per cloth)", "vertex_steps_per_s": vsps}
    # the same cloths with the cloth surrogate network (8x steps, per-mesh fallback): a rejection
    # costs only its own mesh's plain steps
    sur = None
    if args.weights and os.path.exists(args.weights):
        M = len(scene.cloths)
        ref = [sim.get_points(k)[:, 1:4].astype(np.float64) for k in range(M)]
        net = fa.Surrogate(fa.SurrogateConfig(), params=np.load(args.weights, allow_pickle=False))
        ssim = wl.build_gpu(scene)
        ssim.set_surrogate(net)
        ratio = net.config.ratio
        n_big = args.steps // ratio
        if args.warmup:
            ssim.step_surrogate(max(1, args.warmup // ratio))
            for k, rec in enumerate(initial):
                ssim.set_points(k, rec)
        total_ms, mlp_ms, st = ssim.time_surrogate(n_big)
        rel, diverged = [], 0
        for k, c in enumerate(scene.cloths):
            if not (np.isfinite(ref[k]).all() and np.abs(ref[k]).max() < 1e3):
                diverged += 1  # the plain method itself blows up on this cloth (e.g. corner pins)
                continue
            d = ssim.get_points(k)[:, 1:4].astype(np.float64) - ref[k]
            rel.append(float(np.sqrt(np.mean(np.sum(d * d, axis=1)))) / ((max(c.rows, c.cols) - 1) * c.spacing))
        rel = np.array(rel)
        rel = np.where(np.isfinite(rel), rel, np.inf)
        sim_s = n_big * ratio * dt / (total_ms * 1e-3)
        launches = n_big - st.plain_only
        sur = {"sim_s_per_s": sim_s, "speedup_vs_plain": sim_s / (args.steps * dt / wall), "big_steps": n_big,
               "fallbacks": st.fallbacks, "mesh_fallbacks": st.mesh_fallbacks,
               "accepted_mesh_steps_frac": 1.0 - st.mesh_fallbacks / float(M * n_big),
               "plain_only": st.plain_only, "network_launches": launches,
               "rms_error_over_extent_median": float(np.median(rel)) if rel.size else None,
               "rms_error_over_extent_p90": float(np.percentile(rel, 90)) if rel.size else None,
               "cloths_within_1e-3": int(np.sum(rel <= 1e-3)), "cloths_compared": int(rel.size),
               "cloths_diverging_in_plain_method": diverged, "cloths": M,
               "error_definition": "per cloth: RMS of final-position differences vs the plain run after the same "
                                   "simulated time, over the cloth's longer side",
               "mlp_kernel_avg_us": mlp_ms * 1e3 / max(1, launches)}
    line = {
        "metric": "sim-s/s, batched cloths (4096 meshes, per-mesh Config); %HBM peak",
        "value": args.steps * dt / wall,
        "unit": "sim-s/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain springy step, {len(scene.cloths)} grid cloths, sides log-uniform in [8, 256] "
                               f"({N} vertices), per-cloth stiffness/pins/wind, RK4 dt=1/600 s, "
                               f"{args.chunk}-step HIP graphs", "vertices": N, "math": args.math,
                   "tile_shape_per_size_class": chosen},
        "vertex_steps_per_s": args.steps * N / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel_avg_us": kernel_avg_s * 1e6,
                     "bytes_per_launch": BYTES_PER_VERTEX_STEP * N},
        "cpu_baseline": cpu,
        "surrogate": sur,
    }
    print(json.dumps(line), flush=True)


def bench_surfaces(args, np, fa, wl):
    """64 icospheres of uneven size (levels 4-7, 2,145,408 vertices) on the patch kernel."""
    scene = wl.config4()
    sim = wl.build_surface_gpu(scene)
    sim.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT if args.math == "exact" else fa.FP_MATH_FAST)
    N = sum(m.positions.shape[0] for m in scene.meshes)
    S = sum(sim.mesh_counts(k)[1] for k in range(len(scene.meshes)))
    F = sum(sim.mesh_counts(k)[2] for k in range(len(scene.meshes)))
    dt = scene.config.dt
    initial = [sim.get_points(k) for k in range(len(scene.meshes))]
    if args.warmup:
        sim.step_graph(args.warmup, args.chunk)
        for k, rec in enumerate(initial):
            sim.set_points(k, rec)
    sim.synchronize()
    t0 = time.perf_counter()
    kernel_ms = sim.time_steps_graph(args.steps, args.chunk)
    sim.synchronize()
    wall = time.perf_counter() - t0
    # algorithmic bytes per step of a general mesh: the state read and written once (48 B/vertex),
    # a row pointer (4 B/vertex), per directed strut entry the neighbour id and rest length (8 B),
    # per face its 3 vertex ids and one constant (16 B)
    bytes_step = 48 * N + 4 * N + 8 * 2 * S + 16 * F
    kernel_avg_s = kernel_ms * 1e-3 / args.steps
    achieved = bytes_step / kernel_avg_s / 1e9
    cpu = None
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        sub = wl.config4(n_meshes=8, levels=[5, 5, 4, 4, 4, 4, 4, 4])
        o = orc.build_surface_scene(sub)
        n_sub = sum(m.positions.shape[0] for m in sub.meshes)
        t = time.perf_counter()
        o.step(args.cpu_steps)
        cpu_wall = time.perf_counter() - t
        vsps = args.cpu_steps * n_sub / cpu_wall
        cpu = {"value": vsps / N * dt, "unit": "sim-s/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps of 8 icospheres ({n_sub} vertices), scaled by vertex count to "
                         f"{N} vertices", "vertex_steps_per_s": vsps}
    sur = None
    wpath = os.path.join(ROOT, "feriphys_amd", "weights", "surface_surrogate.npy")
    if args.weights and os.path.exists(wpath):
        # the same scene with the irregular-mesh surrogate (8x steps, per-mesh fallback), against the
        # plain run's final positions (the simulation above ran exactly args.steps plain steps)
        ref = [sim.get_points(k)[:, 1:4].astype(np.float64) for k in range(len(scene.meshes))]
        net = fa.Surrogate(fa.SurrogateConfig(), params=np.load(wpath, allow_pickle=False))
        ssim = wl.build_surface_gpu(scene)
        ssim.set_surrogate(net)
        n_big = args.steps // net.config.ratio
        if args.warmup:
            ssim.step_surrogate(max(1, args.warmup // net.config.ratio))
            for k, rec in enumerate(initial):
                ssim.set_points(k, rec)
        total_ms, mlp_ms, st = ssim.time_surrogate(n_big)
        err2, ext = 0.0, max(float(np.ptp(m.positions, axis=0).max()) for m in scene.meshes)
        for k in range(len(scene.meshes)):
            d = ssim.get_points(k)[:, 1:4].astype(np.float64) - ref[k]
            err2 += float(np.sum(d * d))
        launches = n_big - st.plain_only
        sur = {"sim_s_per_s": n_big * net.config.ratio * dt / (total_ms * 1e-3),
               "speedup_vs_plain": (n_big * net.config.ratio * dt / (total_ms * 1e-3)) / (args.steps * dt / wall),
               "big_steps": n_big, "fallbacks": st.fallbacks, "mesh_fallbacks": st.mesh_fallbacks,
               "plain_only": st.plain_only, "network_launches": launches,
               "rms_error_over_extent": float(np.sqrt(err2 / N)) / ext,
               "error_definition": "final positions vs the plain run after the same simulated time; extent = "
                                   f"the largest mesh's diameter {ext:.3f} m",
               "mlp_kernel_avg_us": mlp_ms * 1e3 / max(1, launches)}
    line = {
        "metric": "sim-s/s, irregular meshes (64 icospheres); %HBM peak",
        "value": args.steps * dt / wall, "unit": "sim-s/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain springy step, 64 radially perturbed icospheres (levels 4-7), {N} vertices, "
                               f"{S} struts, {F} faces, ground plane, RK4 dt=1/600 s", "math": args.math},
        "vertex_steps_per_s": args.steps * N / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC_SURFACES.get(args.math),
                     "kernel_avg_us": kernel_avg_s * 1e6, "bytes_per_launch": bytes_step},
        "cpu_baseline": cpu,
        "surrogate": sur,
    }
    print(json.dumps(line), flush=True)


def rigid_scene(np, n, seed=0):
    """n unit cubes on a jittered lattice 1-40 m above a ground plane, masses 0.5-4 kg, random
    impulses; restitution 0 (the reference's vertex impulses add energy in resting contact above
    0), gravity 9.81 m/s^2 as a force, RK4 at dt = 1/600 s."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    i = np.arange(n)
    pos = np.stack([(i % side) * 2.0 - side, 1.0 + rng.uniform(0.0, 39.0, n), (i // side) * 2.0 - side], 1)
    mass = rng.uniform(0.5, 4.0, n)
    imp = rng.normal(scale=2.0, size=(n, 3))
    at = rng.uniform(-0.5, 0.5, size=(n, 3))
    half = float(side) + 2.0
    floor = (np.array([[-half, 0, half], [half, 0, half], [half, 0, -half], [-half, 0, -half]], np.float32),
             np.array([0, 1, 2, 0, 2, 3], np.uint64))
    return pos.astype(np.float32), mass.astype(np.float32), imp.astype(np.float32), at.astype(np.float32), floor


def bench_rigid(args, np, fa, wl):
    """1,048,576 unit cubes (the reference's RigidBody) tumbling onto a ground plane, one GPU thread
    per body.  A "step" is one rigidbody::Simulation::step of every body."""
    n = 1 << 20
    pos, mass, imp, at, floor = rigid_scene(np, n)
    bodies = [fa.RigidBody(pos[i], float(mass[i])) for i in range(n)]
    sim = fa.RigidSimulation(bodies, [fa.CollidableMesh(*floor)])
    rec = sim.states()
    rec[:, 7:10] = rec[:, 7:10] + imp                  # apply_impulse: p += J, L += r x J
    rec[:, 10:13] = rec[:, 10:13] + np.cross(at, imp)
    sim.set_states(rec)
    dt = float(np.float32(1.0 / 600.0))
    sim.config = fa.RigidConfig(fa.Integration.Rk4, dt, 0.0, (0.0, -9.81, 0.0), (0.0, 0.0, 0.0))
    if args.warmup:
        sim.step(args.warmup)
        sim.set_states(rec)
    t0 = time.perf_counter()
    kernel_ms = sim.time_steps(args.steps)
    wall = time.perf_counter() - t0
    out = sim.states()
    if not np.isfinite(out).all():
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)
    kernel_avg_s = kernel_ms * 1e-3 / args.steps
    bytes_step = 2 * 29 * 4 * n  # the 29-float record read and written once per body-step
    achieved = bytes_step / kernel_avg_s / 1e9
    cpu = None
    if args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as orc

        m = 4096
        o = orc.RigidSimulation(pos[:m], mass[:m], [orc.Obstacle(*floor)])
        o.set_states(rec[:m])
        o.set_config(integration=1, dt=dt, restitution=0.0, gravity=(0.0, -9.81, 0.0))
        steps = 10 * args.cpu_steps
        t = time.perf_counter()
        o.step(steps)
        cpu_wall = time.perf_counter() - t
        bsps = steps * m / cpu_wall
        cpu = {"value": bsps / n * dt, "unit": "sim-s/s", "cores": 1, "kind": "port",
               "sample": f"{steps} steps of the first {m} bodies, scaled by body count to {n}",
               "body_steps_per_s": bsps}
    line = {
        "metric": "sim-s/s, 1M rigid cubes; %HBM peak",
        "value": args.steps * dt / wall, "unit": "sim-s/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"rigidbody::Simulation::step of {n} unit cubes (random impulses, restitution 0) "
                               f"onto a ground plane, RK4 dt=1/600 s", "bodies": n},
        "body_steps_per_s": args.steps * n / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel_avg_us": kernel_avg_s * 1e6,
                     "bytes_per_launch": bytes_step},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
