"""Benchmark: the springy step of a 1024 x 1024 (1,048,576-vertex) grid cloth draping over a
sphere collider on one MI355X.

One "step" is one Simulation::step of the whole cloth (forces, RK4, collision, pins) at
dt = 1/600 s.  The cloth's state is resident in HBM before the timed region; the timed region is
exactly K steps bracketed by device synchronisations, replayed from captured HIP graphs.

Both math forms of the step are timed: EXACT (the reference's f32 results bit for bit) and FAST
(one-step forces and states within 1e-5 relative of the reference; the parity tests hold it to
that).  The headline is the faster form that meets its bar, or the learned surrogate when its
600-step run stays within 1e-3 of the exact plain run and is faster still.

Prints ONE JSON line:
  value           simulated seconds per wall second (K * dt / wall time) of the headline form
  roofline        the headline's step kernel against HBM bandwidth: 48 algorithmic bytes per
                  vertex-step (x, y, z, vx, vy, vz read once and written once, f32) over the
                  kernel's average duration measured with HIP events on the simulation's stream
  surrogate       the learned big step over the full 600-step horizon (75 big steps of 8 dt),
                  compared with the exact plain run after the same number of dt
  batched,        the plain rates of the 4,096-cloth batch, the 64-icosphere workload and the
  surfaces,       1,048,576 rigid cubes, with their rooflines (omitted with --no-extra)
  rigid
  cpu_baseline    the oracle (a single-threaded C++ restatement of the reference step) on a bounded
                  sample of the same workload, on this host's cores
"""
import argparse
import hashlib
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_VERTEX_STEP = 48
SURROGATE_HORIZON = 600  # plain steps the surrogate leg covers, whatever --steps is
CSRC = os.path.join(ROOT, "feriphys_amd", "csrc")
COMMON = ["step_device.hpp", "vec.hpp", "exact_ops.hpp", "fdlibm_atan2f.hpp", "kernels.hpp"]
# the source files each measured kernel is built from; a counter reading below is reported only
# while these files still hash to the value it was measured on
KERNEL_SOURCES = {
    # sim.cpp deals the tiles to the XCDs, which sets how often halo rows are fetched twice;
    # obstacles.cpp builds the broad-phase grids the collision search walks
    "grid_tiled_kernel<ExactMath,16,16>": ["grid_tiled.hip", "obstacles.cpp", "sim.cpp"] + COMMON,
    "grid_fast_kernel<15,15>": ["grid_fast.hip", "grid_fast_core.hpp", "grid_fast.hpp", "obstacles.cpp", "sim.cpp"] + COMMON,
    "patch_kernel<ExactMath>": ["kernels.hip"] + COMMON,
}
# rocprofv3 readings per dispatch on the 1024 x 1024 cloth (first steps) and the 64 icospheres:
# FETCH_SIZE and WRITE_SIZE in KiB, each scaled by what the same counter reads for the calibration
# copy (bench.py --only copy: 2.000 and 1.000); SQ_INSTS_VALU and SQ_WAVES summed over the chip and
# GRBM_GUI_ACTIVE summed over the 8 XCDs (the commands are in the profiles named)
MEASURED = {
    ("grid_tiled_kernel<ExactMath,16,16>", "cloth1024"): {
        "src_sha": "326562afa3f32bc6", "fetch_kib": 15887, "write_kib": 24576,
        "valu_insts": 3.5014e7, "waves": 16384, "grbm_gui_active": 1.0991e6,
        "profiles": ["profiles/grid_tiled_exact_cloth1024.md", "profiles/grid_tiled_cloth1024_pmc.md"]},
    ("grid_fast_kernel<15,15>", "cloth1024"): {
        "src_sha": "765ed741246d1deb", "fetch_kib": 16221, "write_kib": 24734,
        "valu_insts": 1.1775e7, "waves": 19044, "grbm_gui_active": 5.3408e5,
        "profiles": ["profiles/grid_fast_cloth1024.md"]},
    ("patch_kernel<ExactMath>", "icospheres64"): {
        "src_sha": "267a4afff89029a1", "fetch_kib": 216816, "write_kib": 50338,
        "profiles": ["profiles/batched_and_surfaces.md"]},
}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
SIMDS, XCDS = 1024, 8


def src_sha(kernel):
    h = hashlib.sha256()
    for f in KERNEL_SOURCES[kernel]:
        with open(os.path.join(CSRC, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured(kernel, workload):
    """The counter readings for (kernel, workload) if they were taken on the current sources."""
    m = MEASURED.get((kernel, workload))
    if m is None or kernel not in KERNEL_SOURCES or src_sha(kernel) != m["src_sha"]:
        return None
    return m


def traffic_of(kernel, workload):
    m = measured(kernel, workload)
    return None if m is None else (m["fetch_kib"] * FETCH_SCALE + m["write_kib"] * WRITE_SCALE) * 1024


def valu_of(kernel, workload):
    """VALU issue of a kernel: a wave64 f32 instruction holds a 16-lane SIMD for 4 cycles."""
    m = measured(kernel, workload)
    if m is None or "valu_insts" not in m:
        return None
    issue = m["valu_insts"] * 4 / SIMDS
    active = m["grbm_gui_active"] / XCDS
    return {"insts_per_wave": m["valu_insts"] / m["waves"], "issue_cycles_per_simd": issue,
            "active_cycles": active, "issue_frac": issue / active, "profiles": m["profiles"]}


def graph_chunk(steps, chunk):
    """Steps per captured graph: at most `chunk` and at most `steps`, even (so one graph serves
    both state-buffer parities' starting point), at least 2; 0 (eager launches) when `chunk` is 0."""
    if chunk <= 0:
        return 0
    c = max(2, min(chunk, steps))
    return c - (c % 2)


def launch_desc(chunk):
    return f"{chunk}-step HIP graphs" if chunk > 0 else "eager launches"


def time_plain(sim, initial_per_mesh, steps, warmup, chunk):
    """Times `steps` steps replayed from `chunk`-step HIP graphs, starting from the given records:
    an untimed warm-up of at least one chunk (which captures the graph), a restore of the initial
    records (which keeps the captured graph), then the timed steps, bracketed by synchronisations;
    HIP events on the simulation's stream measure the device time of the same steps."""
    def restore():
        for k, rec in enumerate(initial_per_mesh):
            sim.set_points(k, rec)
    restore()
    if chunk > 0:
        w = max(warmup, chunk)
        sim.step_graph(-(-w // chunk) * chunk, chunk)
    elif warmup:
        sim.step(warmup)
    restore()
    sim.synchronize()
    t0 = time.perf_counter()
    kernel_ms = sim.time_steps_graph(steps, chunk) if chunk > 0 else sim.time_steps(steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    return wall, kernel_ms


def plain_leg(sim, fa, kern, math, initial, steps, warmup, chunk, N, dt, np):
    sim.set_kernel(kern, fa.FP_MATH_EXACT if math == "exact" else fa.FP_MATH_FAST)
    wall, kernel_ms = time_plain(sim, [initial], steps, warmup, chunk)
    pts = sim.get_points(0)
    kavg = kernel_ms * 1e-3 / steps
    return {"math": math, "sim_s_per_s": steps * dt / wall, "ms_per_step": wall * 1e3 / steps,
            "vertex_steps_per_s": steps * N / wall, "kernel_avg_us": kavg * 1e6,
            "achieved_gbs": BYTES_PER_VERTEX_STEP * N / kavg / 1e9,
            "hbm_frac": BYTES_PER_VERTEX_STEP * N / kavg / 1e9 / HBM_PEAK_GBS,
            "finite": bool(np.isfinite(pts).all())}, pts


def run_from(sim, fa, kern, math, initial, steps, chunk):
    """Positions (f64) after `steps` plain steps from `initial` under the given math form."""
    import numpy as np
    sim.set_kernel(kern, fa.FP_MATH_EXACT if math == "exact" else fa.FP_MATH_FAST)
    sim.set_points(0, initial)
    if chunk > 0:
        sim.step_graph(steps, chunk)
    else:
        sim.step(steps)
    return sim.get_points(0)[:, 1:4].astype(np.float64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--n", type=int, default=1024, help="cloth side (vertices)")
    ap.add_argument("--math", choices=["auto", "exact", "fast"], default="auto",
                    help="auto: time both forms and headline the faster")
    ap.add_argument("--kernel", choices=["auto", "grid", "grid_simple", "csr"], default="auto")
    ap.add_argument("--tile", choices=["default", "15x15", "7x31", "16x16"], default="default")
    ap.add_argument("--cpu-steps", type=int, default=5, help="oracle steps for the CPU baseline (0: skip; 5 on the 1M cloth take about 10 s of one core)")
    ap.add_argument("--workload", choices=["cloth", "batched", "surfaces", "rigid"], default="cloth",
                    help="cloth: one 1024x1024 cloth + sphere (default); batched: 4096 cloths of uneven sizes; "
                         "surfaces: 64 perturbed icospheres (2.1M vertices) on a ground plane; rigid: 1M unit "
                         "cubes tumbling onto a ground plane")
    ap.add_argument("--only", choices=["all", "plain", "mlp", "copy"], default="all",
                    help="plain: only the timed plain steps of the cloth (for profiling the step kernel); "
                         "mlp: only the surrogate network kernel over the cloth's vertices, --steps launches; "
                         "copy: --steps device copies of 256 MiB (the calibration of the HBM byte counters)")
    ap.add_argument("--no-extra", action="store_true", help="leave the batched, surfaces and rigid keys out of the line")
    ap.add_argument("--chunk", type=int, default=100, help="steps per captured HIP graph")
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
    if args.only == "mlp":
        return only_mlp(args, np, fa, wl)
    if args.only == "copy":
        return only_copy(args)

    scene = wl.config1(n=args.n, steps=args.steps)
    sim = wl.build_gpu(scene)
    kern = {"auto": fa.FP_KERNEL_AUTO, "grid": fa.FP_KERNEL_GRID, "grid_simple": fa.FP_KERNEL_GRID_SIMPLE,
            "csr": fa.FP_KERNEL_CSR}[args.kernel]
    if args.tile != "default":
        sim.set_tile_shape(0, {"15x15": 0, "7x31": 1, "16x16": 2}[args.tile])
    N = scene.cloths[0].rows * scene.cloths[0].cols
    dt = scene.config.dt
    chunk = graph_chunk(args.steps, args.chunk)
    initial = sim.get_points(0)

    maths = ["exact", "fast"] if args.math == "auto" else [args.math]
    legs, finals = {}, {}
    for m in maths:
        legs[m], finals[m] = plain_leg(sim, fa, kern, m, initial, args.steps, args.warmup, chunk, N, dt, np)
        print(f"[bench] plain {m}: {legs[m]['kernel_avg_us']:.1f} us/step (HIP events), "
              f"{legs[m]['sim_s_per_s']:.2f} sim-s/s", file=sys.stderr, flush=True)
    if not all(l["finite"] for l in legs.values()):
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)
    best = max(legs, key=lambda m: legs[m]["sim_s_per_s"])
    kname = {"exact": "grid_tiled_kernel<ExactMath,16,16>", "fast": "grid_fast_kernel<15,15>"}
    for m in legs:
        legs[m]["kernel"] = kname[m]
        legs[m]["traffic"] = traffic_of(kname[m], "cloth1024") if args.n == 1024 else None
        legs[m]["valu_issue"] = valu_of(kname[m], "cloth1024") if args.n == 1024 else None

    if args.only == "plain":
        print(json.dumps(headline_line(args, legs, best, None, None, {}, N, dt, scene)), flush=True)
        return

    # the two forms over the surrogate's 600-step horizon: the exact run is what the surrogate is measured against
    horizon_chunk = graph_chunk(SURROGATE_HORIZON, args.chunk)
    ref_exact = run_from(sim, fa, kern, "exact", initial, SURROGATE_HORIZON, horizon_chunk)
    compare = {"steps": SURROGATE_HORIZON}
    if "fast" in legs:
        xf = run_from(sim, fa, kern, "fast", initial, SURROGATE_HORIZON, horizon_chunk)
        d = np.linalg.norm(xf - ref_exact, axis=1)
        compare.update({"fast_vs_exact_max_m": float(d.max()),
                        "fast_vs_exact_rms_over_extent": float(np.sqrt(np.mean(d * d))) / ((args.n - 1) * 0.05)})

    sur = surrogate_leg(args, np, fa, wl, sim, kern, ref_exact, initial, N, dt, legs)

    # PCIe-inclusive rate (never the headline): a caller that reads the 10-float records back after
    # every step, as the reference's demo renders every frame, into one pinned host buffer it
    # reuses (as a render loop keeps its staging buffer); host clock over 20 frames
    sim.set_surrogate(None)
    sim.set_kernel(kern, fa.FP_MATH_EXACT if best == "exact" else fa.FP_MATH_FAST)
    sim.set_points(0, initial)
    host = fa.pinned_empty((N, 10))
    sim.get_points(0, out=host)
    sim.synchronize()
    frames = 20
    tr = time.perf_counter()
    for _ in range(frames):
        sim.step(1)
        sim.get_points(0, out=host)
    frame_s = (time.perf_counter() - tr) / frames
    with_readback = {"ms_per_frame": frame_s * 1e3, "sim_s_per_s": dt / frame_s,
                     "readback_bytes_per_frame": 40 * N, "readback_gb_per_s": 40 * N / frame_s / 1e9,
                     "host_buffer": "pinned, reused", "frames": frames, "math": best}
    del sim

    extra = {}
    if not args.no_extra:
        extra["batched"] = batched_plain(args, np, fa, wl)
        extra["surfaces"] = surfaces_plain(args, np, fa, wl)
        extra["rigid"], _, _, _ = rigid_plain(args, np, fa)

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

    line = headline_line(args, legs, best, sur, compare, extra, N, dt, scene)
    line["with_readback"] = with_readback
    line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def headline_line(args, legs, best, sur, compare, extra, N, dt, scene):
    """The JSON line: the faster plain form, or the surrogate when it meets its bar and is faster."""
    use_sur = sur is not None and sur["meets_bar"] and sur["sim_s_per_s"] > legs[best]["sim_s_per_s"]
    L = legs[best]
    roof = {"bound": "hbm", "achieved": L["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": L["hbm_frac"], "traffic": L["traffic"], "kernel": L["kernel"],
            "kernel_avg_us": L["kernel_avg_us"], "bytes_per_launch": BYTES_PER_VERTEX_STEP * N}
    if use_sur:
        r = sur["mlp_roofline"]
        roof = {"bound": "mfma", "achieved": r["achieved"], "peak": r["peak"], "unit": "TFLOP/s", "frac": r["frac"],
                "traffic": None, "kernel": "sur_mlp_kernel", "kernel_avg_us": r["kernel_avg_us"],
                "flop_per_launch": r["flop_per_launch"]}
    value = sur["sim_s_per_s"] if use_sur else L["sim_s_per_s"]
    line = {
        "metric": "sim-s/s, 1M-vert cloth: surrogate vs plain @<=1e-3 pos err; %HBM peak",
        "value": value,
        "value_source": "surrogate" if use_sur else f"plain ({best} math)",
        "unit": "sim-s/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / value * 1e3,  # wall time per dt of simulated time
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain springy step, {args.n}x{args.n} grid cloth ({N} vertices), top edge pinned, "
                               f"320-face icosphere collider, RK4 dt=1/600 s, {launch_desc(graph_chunk(args.steps, args.chunk))}",
                   "vertices": N, "dt": dt, "kernel": args.kernel, "math": best,
                   "math_bars": {"exact": "bit-identical to the reference's f32 step",
                                 "fast": "one-step forces and states within 1e-5 relative of the reference"}},
        "vertex_steps_per_s": N * value / dt,
        "plain": legs,
        "roofline": roof,
        "valu_issue": L["valu_issue"],
    }
    if compare is not None:
        line["plain_600_step_comparison"] = compare
    if sur is not None:
        line["surrogate"] = sur
    line.update(extra)
    return line


MLP_FLOP_PER_VERTEX = 2 * (72 * 128 + 128 * 128 + 128 * 128 + 128 * 3)  # 84,736
MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_leg(args, np, fa, wl, sim, kern, ref_exact, initial, N, dt, legs):
    """The same cloth stepped by the learned surrogate (8x steps, per-mesh fallback on) over the
    full 600-step horizon whatever --steps is, timed on the host clock around the whole run (it
    reads each big step's diagnostics back); the network kernel's launches timed with HIP events.
    Error: the final positions against the exact plain run after the same number of dt (the plain
    step of springy/simulation.rs:27-45 is the reference for the error)."""
    if not args.weights or not os.path.exists(args.weights):
        return None
    params = np.load(args.weights, allow_pickle=False)
    net = fa.Surrogate(fa.SurrogateConfig(), params=params)
    ratio = net.config.ratio
    n_big = SURROGATE_HORIZON // ratio
    assert n_big * ratio == SURROGATE_HORIZON, "the surrogate must cover exactly the plain run's dt count"
    sim.set_kernel(kern, fa.FP_MATH_FAST if "fast" in legs else fa.FP_MATH_EXACT)
    sim.set_surrogate(net)
    sim.set_surrogate_shape(args.sur_shape)
    sim.set_points(0, initial)
    sim.step_surrogate(1)  # warm-up (first launches), then the initial state again
    sim.set_points(0, initial)
    sim.set_surrogate(net)  # re-attaching clears the fallback backoff the warm-up may have left
    sim.set_surrogate_shape(args.sur_shape)
    sim.synchronize()
    total_ms, mlp_ms, st = sim.time_surrogate(n_big)
    assert st.big_steps == n_big
    x = sim.get_points(0)[:, 1:4].astype(np.float64)
    x0 = initial[:, 1:4].astype(np.float64)
    extent = (args.n - 1) * 0.05
    e = np.linalg.norm(x - ref_exact, axis=1)
    disp = float(np.sqrt(np.mean(np.sum((ref_exact - x0) ** 2, axis=1))))
    rms = float(np.sqrt(np.mean(e * e)))
    sim_s = n_big * ratio * dt / (total_ms * 1e-3)
    launches = n_big - st.plain_only  # big steps that ran the network
    mlp_avg_s = mlp_ms * 1e-3 / max(1, launches)
    tflops = MLP_FLOP_PER_VERTEX * N / mlp_avg_s / 1e12 if mlp_avg_s > 0 and launches > 0 else None
    finite = bool(np.isfinite(x).all())
    return {
        "sim_s_per_s": sim_s,
        "speedup_vs_plain": {m: sim_s / legs[m]["sim_s_per_s"] for m in legs},
        "big_steps": n_big, "ratio": ratio, "plain_steps_covered": n_big * ratio,
        "ms_per_big_step": total_ms / max(1, n_big),
        "rms_error_over_extent": rms / extent,
        "max_error_over_extent": float(e.max()) / extent,
        "rms_error_over_rms_displacement": rms / disp if disp > 0 else None,
        "rms_displacement_m": disp,
        "meets_bar": finite and rms / extent <= 1e-3,
        "finite": finite,
        "fallbacks": st.fallbacks, "plain_only": st.plain_only, "network_launches": launches,
        "max_strain_residual": st.max_strain_residual,
        "max_energy_residual": st.max_energy_residual,
        "error_definition": f"RMS over vertices of |x_surrogate - x_plain| after {n_big * ratio} dt, the plain run "
                            f"being the exact step from the same initial state; extent = cloth width {extent:.2f} m",
        "mlp_roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tflops / MFMA_BF16_PEAK_TFLOPS if tflops else None,
                         "kernel_avg_us": mlp_avg_s * 1e6, "flop_per_launch": MLP_FLOP_PER_VERTEX * N},
    }


def only_mlp(args, np, fa, wl):
    """The surrogate network kernel alone over the 1024 x 1024 cloth's vertices (features gathered
    from the state, three GELU layers, the output), --steps launches; for rocprofv3."""
    scene = wl.config1(n=args.n, steps=args.steps)
    sim = wl.build_gpu(scene)
    net = fa.Surrogate(fa.SurrogateConfig(), params=np.load(args.weights, allow_pickle=False))
    sim.set_surrogate(net)
    sim.set_surrogate_shape(args.sur_shape)
    N = scene.cloths[0].rows * scene.cloths[0].cols
    for _ in range(max(1, args.warmup // 10)):
        sim.surrogate_accel(0)
    t = time.perf_counter()
    for _ in range(args.steps):
        a = sim.surrogate_accel(0)
    wall = time.perf_counter() - t
    print(json.dumps({"metric": "surrogate network launches (profiling entry point)", "value": args.steps / wall,
                      "unit": "launches/s (host clock, incl. readback)", "launches": args.steps, "vertices": N,
                      "finite": bool(np.isfinite(a).all())}), flush=True)


def only_copy(args):
    """A plain device copy of 256 MiB, --steps times: what FETCH_SIZE and WRITE_SIZE read for a
    kernel that moves a known number of bytes (each reads 256 MiB once and writes it once)."""
    import torch

    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    print(json.dumps({"metric": "device copy (counter calibration entry point)", "bytes_read_per_copy": 4 * n,
                      "bytes_written_per_copy": 4 * n, "copies": args.steps,
                      "value": 2 * 4 * n * args.steps / wall / 1e9, "unit": "GB/s (host clock)"}), flush=True)


def plain_rates(sim, fa, initial, steps, warmup, chunk, n_vertices, dt, bytes_step):
    out = {}
    for m in ("exact", "fast"):
        sim.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT if m == "exact" else fa.FP_MATH_FAST)
        wall, kernel_ms = time_plain(sim, initial, steps, warmup, chunk)
        kavg = kernel_ms * 1e-3 / steps
        out[m] = {"sim_s_per_s": steps * dt / wall, "ms_per_step": wall * 1e3 / steps,
                  "vertex_steps_per_s": steps * n_vertices / wall, "kernel_avg_us": kavg * 1e6,
                  "achieved_gbs": bytes_step / kavg / 1e9, "hbm_frac": bytes_step / kavg / 1e9 / HBM_PEAK_GBS}
    return out


def batched_plain(args, np, fa, wl):
    """Plain rates of the 4,096-cloth batch (per-cloth Config, tile shapes chosen by the autotuner)."""
    scene = wl.config2(steps=args.steps)
    sim = wl.build_gpu(scene)
    N = sum(c.rows * c.cols for c in scene.cloths)
    chosen = sim.autotune(10)
    initial = [sim.get_points(k) for k in range(len(scene.cloths))]
    chunk = graph_chunk(args.steps, args.chunk)
    rates = plain_rates(sim, fa, initial, args.steps, args.warmup, chunk, N, scene.config.dt, BYTES_PER_VERTEX_STEP * N)
    return {"workload": f"{len(scene.cloths)} grid cloths, sides log-uniform in [8, 256] ({N} vertices), per-cloth "
                        f"stiffness/pins/wind, RK4 dt=1/600 s, {launch_desc(chunk)}",
            "vertices": N, "tile_shape_per_size_class": chosen, "steps": args.steps,
            "bytes_per_vertex_step": BYTES_PER_VERTEX_STEP, "plain": rates}


def surface_bytes(N, S, F):
    # algorithmic bytes per step of a general mesh: the state read and written once (48 B/vertex),
    # a row pointer (4 B/vertex), per directed strut entry the neighbour id and rest length (8 B),
    # per face its 3 vertex ids and one constant (16 B)
    return 48 * N + 4 * N + 8 * 2 * S + 16 * F


def surfaces_plain(args, np, fa, wl):
    """Plain rates of the 64-icosphere workload (patch kernel)."""
    scene = wl.config4()
    sim = wl.build_surface_gpu(scene)
    N = sum(m.positions.shape[0] for m in scene.meshes)
    S = sum(sim.mesh_counts(k)[1] for k in range(len(scene.meshes)))
    F = sum(sim.mesh_counts(k)[2] for k in range(len(scene.meshes)))
    initial = [sim.get_points(k) for k in range(len(scene.meshes))]
    chunk = graph_chunk(args.steps, args.chunk)
    bstep = surface_bytes(N, S, F)
    rates = plain_rates(sim, fa, initial, args.steps, args.warmup, chunk, N, scene.config.dt, bstep)
    rates["exact"]["traffic"] = traffic_of("patch_kernel<ExactMath>", "icospheres64")
    return {"workload": f"64 radially perturbed icospheres (levels 4-7), {N} vertices, {S} struts, {F} faces, ground "
                        f"plane, RK4 dt=1/600 s, {launch_desc(chunk)}",
            "vertices": N, "struts": S, "faces": F, "steps": args.steps, "bytes_per_step": bstep, "plain": rates}


def bench_batched(args, np, fa, wl):
    """4096 cloths (sides log-uniform in [8, 256], ~19.7M vertices), each with its own Config, one
    tiled launch per tile shape per step, steps replayed from HIP graphs of `chunk` steps; then the
    cloth surrogate on the same batch."""
    scene = wl.config2(steps=args.steps)
    sim = wl.build_gpu(scene)
    math = "exact" if args.math != "fast" else "fast"
    sim.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT if math == "exact" else fa.FP_MATH_FAST)
    N = sum(c.rows * c.cols for c in scene.cloths)
    dt = scene.config.dt
    chosen = sim.autotune(10)
    initial = [sim.get_points(k) for k in range(len(scene.cloths))]
    chunk = graph_chunk(args.steps, args.chunk)
    wall, kernel_ms = time_plain(sim, initial, args.steps, args.warmup, chunk)
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
    # costs only its own mesh's plain steps; compared with the plain run over the same dt count
    sur = None
    if args.weights and os.path.exists(args.weights):
        M = len(scene.cloths)
        net = fa.Surrogate(fa.SurrogateConfig(), params=np.load(args.weights, allow_pickle=False))
        ratio = net.config.ratio
        n_big = SURROGATE_HORIZON // ratio
        for k, rec in enumerate(initial):
            sim.set_points(k, rec)
        sim.step_graph(n_big * ratio, graph_chunk(n_big * ratio, args.chunk))
        ref = [sim.get_points(k)[:, 1:4].astype(np.float64) for k in range(M)]
        for k, rec in enumerate(initial):
            sim.set_points(k, rec)
        sim.set_surrogate(net)
        total_ms, mlp_ms, st = sim.time_surrogate(n_big)
        rel, diverged = [], 0
        for k, c in enumerate(scene.cloths):
            if not (np.isfinite(ref[k]).all() and np.abs(ref[k]).max() < 1e3):
                diverged += 1  # the plain method itself blows up on this cloth (e.g. corner pins)
                continue
            d = sim.get_points(k)[:, 1:4].astype(np.float64) - ref[k]
            rel.append(float(np.sqrt(np.mean(np.sum(d * d, axis=1)))) / ((max(c.rows, c.cols) - 1) * c.spacing))
        rel = np.array(rel)
        rel = np.where(np.isfinite(rel), rel, np.inf)
        sim_s = n_big * ratio * dt / (total_ms * 1e-3)
        launches = n_big - st.plain_only
        sur = {"sim_s_per_s": sim_s, "speedup_vs_plain": sim_s / (args.steps * dt / wall), "big_steps": n_big,
               "plain_steps_covered": n_big * ratio,
               "fallbacks": st.fallbacks, "mesh_fallbacks": st.mesh_fallbacks,
               "accepted_mesh_steps_frac": 1.0 - st.mesh_fallbacks / float(M * n_big),
               "plain_only": st.plain_only, "network_launches": launches,
               "rms_error_over_extent_median": float(np.median(rel)) if rel.size else None,
               "rms_error_over_extent_p90": float(np.percentile(rel, 90)) if rel.size else None,
               "cloths_within_1e-3": int(np.sum(rel <= 1e-3)), "cloths_compared": int(rel.size),
               "cloths_diverging_in_plain_method": diverged, "cloths": M,
               "error_definition": "per cloth: RMS of final-position differences vs the plain run after the same "
                                   "number of dt, over the cloth's longer side",
               "mlp_kernel_avg_us": mlp_ms * 1e3 / max(1, launches)}
    line = {
        "metric": "sim-s/s, batched cloths (4096 meshes, per-mesh Config); %HBM peak",
        "value": args.steps * dt / wall,
        "unit": "sim-s/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain springy step, {len(scene.cloths)} grid cloths, sides log-uniform in [8, 256] "
                               f"({N} vertices), per-cloth stiffness/pins/wind, RK4 dt=1/600 s, "
                               f"{launch_desc(chunk)}", "vertices": N, "math": math,
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
    """64 icospheres of uneven size (levels 4-7, 2,145,408 vertices) on the patch kernel; then the
    irregular-mesh surrogate on the same scene."""
    scene = wl.config4()
    sim = wl.build_surface_gpu(scene)
    math = "exact" if args.math != "fast" else "fast"
    sim.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT if math == "exact" else fa.FP_MATH_FAST)
    N = sum(m.positions.shape[0] for m in scene.meshes)
    S = sum(sim.mesh_counts(k)[1] for k in range(len(scene.meshes)))
    F = sum(sim.mesh_counts(k)[2] for k in range(len(scene.meshes)))
    dt = scene.config.dt
    initial = [sim.get_points(k) for k in range(len(scene.meshes))]
    chunk = graph_chunk(args.steps, args.chunk)
    wall, kernel_ms = time_plain(sim, initial, args.steps, args.warmup, chunk)
    bytes_step = surface_bytes(N, S, F)
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
        # the irregular-mesh surrogate (8x steps, per-mesh fallback), against the plain run over the
        # same number of dt
        net = fa.Surrogate(fa.SurrogateConfig(), params=np.load(wpath, allow_pickle=False))
        n_big = SURROGATE_HORIZON // net.config.ratio
        for k, rec in enumerate(initial):
            sim.set_points(k, rec)
        sim.step_graph(n_big * net.config.ratio, graph_chunk(n_big * net.config.ratio, args.chunk))
        ref = [sim.get_points(k)[:, 1:4].astype(np.float64) for k in range(len(scene.meshes))]
        for k, rec in enumerate(initial):
            sim.set_points(k, rec)
        sim.set_surrogate(net)
        total_ms, mlp_ms, st = sim.time_surrogate(n_big)
        err2, ext = 0.0, max(float(np.ptp(m.positions, axis=0).max()) for m in scene.meshes)
        for k in range(len(scene.meshes)):
            d = sim.get_points(k)[:, 1:4].astype(np.float64) - ref[k]
            err2 += float(np.sum(d * d))
        launches = n_big - st.plain_only
        sur = {"sim_s_per_s": n_big * net.config.ratio * dt / (total_ms * 1e-3),
               "speedup_vs_plain": (n_big * net.config.ratio * dt / (total_ms * 1e-3)) / (args.steps * dt / wall),
               "big_steps": n_big, "plain_steps_covered": n_big * net.config.ratio,
               "fallbacks": st.fallbacks, "mesh_fallbacks": st.mesh_fallbacks,
               "plain_only": st.plain_only, "network_launches": launches,
               "rms_error_over_extent": float(np.sqrt(err2 / N)) / ext,
               "error_definition": "final positions vs the plain run after the same number of dt; extent = "
                                   f"the largest mesh's diameter {ext:.3f} m",
               "mlp_kernel_avg_us": mlp_ms * 1e3 / max(1, launches)}
    line = {
        "metric": "sim-s/s, irregular meshes (64 icospheres); %HBM peak",
        "value": args.steps * dt / wall, "unit": "sim-s/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain springy step, 64 radially perturbed icospheres (levels 4-7), {N} vertices, "
                               f"{S} struts, {F} faces, ground plane, RK4 dt=1/600 s, {launch_desc(chunk)}",
                   "math": math},
        "vertex_steps_per_s": args.steps * N / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_of("patch_kernel<ExactMath>", "icospheres64") if math == "exact" else None,
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


RIGID_BODIES = 1 << 20
RIGID_BYTES_PER_BODY_STEP = 2 * 29 * 4  # the 29-float record read and written once per body-step


def rigid_plain(args, np, fa):
    """The 1,048,576-cube rigid step timed over args.steps steps after args.warmup untimed ones (host
    clock for the rate, HIP events for the kernel); returns (rates, scene pieces, records, dt)."""
    n = RIGID_BODIES
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
        sim.set_states(rec)  # returns once the records are on the device
    t0 = time.perf_counter()
    kernel_ms = sim.time_steps(args.steps)
    wall = time.perf_counter() - t0
    out = sim.states()
    kernel_avg_s = kernel_ms * 1e-3 / args.steps
    bytes_step = RIGID_BYTES_PER_BODY_STEP * n
    achieved = bytes_step / kernel_avg_s / 1e9
    rates = {"workload": f"rigidbody::Simulation::step of {n} unit cubes (random impulses, restitution 0) onto a "
                         f"ground plane, RK4 dt=1/600 s", "bodies": n, "steps": args.steps,
             "sim_s_per_s": args.steps * dt / wall, "ms_per_step": wall * 1e3 / args.steps,
             "body_steps_per_s": args.steps * n / wall, "kernel_avg_us": kernel_avg_s * 1e6,
             "bytes_per_launch": bytes_step, "achieved_gbs": achieved, "hbm_frac": achieved / HBM_PEAK_GBS,
             "finite": bool(np.isfinite(out).all())}
    return rates, (pos, mass, floor), rec, dt


def bench_rigid(args, np, fa, wl):
    """1,048,576 unit cubes (the reference's RigidBody) tumbling onto a ground plane, one GPU thread
    per body.  A "step" is one rigidbody::Simulation::step of every body."""
    n = RIGID_BODIES
    rates, (pos, mass, floor), rec, dt = rigid_plain(args, np, fa)
    if not rates["finite"]:
        print("non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)
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
        "value": rates["sim_s_per_s"], "unit": "sim-s/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": rates["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": rates["workload"], "bodies": n},
        "body_steps_per_s": rates["body_steps_per_s"],
        "roofline": {"bound": "hbm", "achieved": rates["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": rates["hbm_frac"], "traffic": None, "kernel_avg_us": rates["kernel_avg_us"],
                     "bytes_per_launch": rates["bytes_per_launch"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
