"""Trains the cloth surrogate network on the GPU and writes its weights.

    python -m feriphys_amd.train_surrogate --out feriphys_amd/weights/cloth_surrogate.npy

Data: the plain solver's trajectories (600 steps at dt = 1/600 s) of seeded training cloths
(workloads.surrogate_training_scene: one pinned edge, random wind; 2,000 with sides in [12, 96]
and a few with sides in [192, 512]) or, with --scene surfaces, icospheres dropped on a ground
plane (workloads.surface_training_scene), sampled every 8 steps, with normal position noise on
the inputs.  Training: Adam on the library's bf16 MFMA kernels, in three
learning-rate phases.  Evaluation: the 32 x 32 edge-pinned cloth (or six config4 icospheres)
stepped 75 big steps with the fallback off, against 600 plain steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import feriphys_amd as fa  # noqa: E402
from feriphys_amd import workloads as wl  # noqa: E402


def evaluate_surfaces(net, big_steps=75):
    """Six icospheres of the config4 recipe (levels 4-5) on the ground plane, 75 big steps with the
    fallback off against 600 plain steps; error relative to the largest mesh's diameter."""
    sc = wl.config4(n_meshes=6, seed=5, levels=[5, 5, 4, 4, 4, 4])
    ref, sur = wl.build_surface_gpu(sc), wl.build_surface_gpu(sc)
    off = fa.Surrogate(fa.SurrogateConfig(strain_limit=1e30, energy_limit=1e30), params=net.get_params())
    sur.set_surrogate(off)
    extent = max(float(np.ptp(m.positions, axis=0).max()) for m in sc.meshes)
    out = []
    for k in range(1, big_steps + 1):
        ref.step(8)
        sur.step_surrogate(1)
        if k in (1, 2, 5, 10, 20, 40, 75):
            d = np.concatenate([sur.get_points(m)[:, 1:4].astype(np.float64) - ref.get_points(m)[:, 1:4]
                                for m in range(len(sc.meshes))])
            out.append((k, float(np.sqrt(np.mean(np.sum(d * d, axis=1))) / extent)))
    return out


def evaluate(net, big_steps=75, seed=21):
    sc = wl.config0(steps=1, seed=seed, rows=32, cols=32, pins="edge")
    ref, sur = wl.build_gpu(sc), wl.build_gpu(sc)
    off = fa.Surrogate(fa.SurrogateConfig(strain_limit=1e30, energy_limit=1e30), params=net.get_params())
    sur.set_surrogate(off)
    extent = 31 * 0.05
    out = []
    for k in range(1, big_steps + 1):
        ref.step(8)
        sur.step_surrogate(1)
        if k in (1, 2, 5, 10, 20, 40, 75):
            d = sur.get_points(0)[:, 1:4].astype(np.float64) - ref.get_points(0)[:, 1:4]
            out.append((k, float(np.sqrt(np.mean(np.sum(d * d, axis=1))) / extent)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cloths", type=int, default=2000)
    ap.add_argument("--smin", type=int, default=12)
    ap.add_argument("--smax", type=int, default=96)
    ap.add_argument("--big", type=int, default=75, help="big steps (of 8 plain steps) per trajectory")
    ap.add_argument("--noise", type=float, default=1e-3)
    ap.add_argument("--steps", type=int, default=9000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--groups", type=int, default=4, help="trajectories are simulated in this many batches")
    ap.add_argument("--large", type=int, default=16,
                    help="extra cloths with sides in [192, 512] (the tension of long hanging cloths)")
    ap.add_argument("--sphere-groups", type=int, default=0,
                    help="of the cloth groups, how many fall onto a sphere collider (contact data)")
    ap.add_argument("--scene", choices=["cloths", "surfaces"], default="cloths",
                    help="cloths: grid cloths (the 1024 x 1024 cloth's network); surfaces: icospheres (config4's)")
    ap.add_argument("--meshes", type=int, default=400, help="icospheres per training scene (--scene surfaces)")
    ap.add_argument("--out", default="feriphys_amd/weights/cloth_surrogate.npy")
    args = ap.parse_args()

    t0 = time.time()
    net = fa.Surrogate(seed=args.seed)
    per = (args.cloths + args.groups - 1) // args.groups
    for g in range(args.groups if args.scene == "cloths" else 0):
        n = min(per, args.cloths - g * per)
        sc = wl.surrogate_training_scene(n_cloths=n, seed=1000 * (args.seed + 1) + g, smin=args.smin, smax=args.smax,
                                         sphere=g < args.sphere_groups)
        sim = wl.build_gpu(sc)
        net.collect(sim, n_big=args.big, noise=args.noise, seed=args.seed * 7919 + g)
        del sim
        print(f"group {g}: {n} cloths, {net.sample_count()} samples so far ({time.time() - t0:.1f} s)", flush=True)
    if args.scene == "surfaces":
        for g in range(args.groups):
            sc = wl.surface_training_scene(n_meshes=args.meshes, seed=1000 * (args.seed + 1) + g)
            sim = wl.build_surface_gpu(sc)
            net.collect(sim, n_big=args.big, noise=args.noise, seed=args.seed * 7919 + g)
            del sim
            print(f"group {g}: {args.meshes} icospheres, {net.sample_count()} samples so far "
                  f"({time.time() - t0:.1f} s)", flush=True)
    if args.large and args.scene == "cloths":
        sc = wl.surrogate_training_scene(n_cloths=args.large, seed=7777 + args.seed, smin=192, smax=512,
                                         sphere=args.sphere_groups > 0)
        sim = wl.build_gpu(sc)
        net.collect(sim, n_big=args.big, noise=args.noise, seed=args.seed * 7919 + 99)
        del sim
        print(f"large cloths: {args.large}, {net.sample_count()} samples so far ({time.time() - t0:.1f} s)",
              flush=True)
    t1 = time.time()
    phases = [(0.45, 1e-3), (0.35, 3e-4), (0.2, 1e-4)]
    log = []
    for i, (frac, lr) in enumerate(phases):
        steps = max(1, int(args.steps * frac))
        losses = net.train(steps, batch=args.batch, lr=lr, seed=args.seed * 31 + i)
        log.append({"steps": steps, "lr": lr, "loss_first": float(losses[:10].mean()),
                    "loss_last": float(losses[-100:].mean())})
        print(f"phase {i}: {steps} steps at lr {lr}: loss {log[-1]['loss_first']:.4f} -> {log[-1]['loss_last']:.4f} "
              f"({time.time() - t1:.1f} s)", flush=True)
    train_s = time.time() - t1
    ev = evaluate(net) if args.scene == "cloths" else evaluate_surfaces(net)
    for k, e in ev:
        print(f"eval big step {k}: rms position error / extent {e:.2e}", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, net.get_params())
    summary = {"samples": net.sample_count(), "collect_s": t1 - t0, "train_s": train_s, "phases": log,
               "eval_rms_over_extent": ev, "args": vars(args)}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
