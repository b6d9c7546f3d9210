"""Throughput benchmark of the MI355X diffuse-react engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]
                    [--workload NAME] [--no-cpu-baseline]

A step is one diffuse-react iteration over every molecule of the workload,
with the molecules resident in HBM.  Prints one JSON line with the metric
(molecule-timesteps/s), the roofline of the dominant kernel measured with HIP
events on the engine's stream, and a CPU baseline: the oracle (a restatement
of the reference's FP64 CPU path, ISAAC64 stream, one core) timed on a bounded
sample of the same kind of work on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
# algorithmic bytes per volume molecule-timestep: SoA state read once and
# written once (positions + species/flags word + subpartition word)
B_ALG = {"fp64": 64.0, "fp32": 40.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="both", choices=["fp64", "fp32", "both"],
                    help="both: FP32+FP64-fallback is the reported value, the plain FP64 form runs beside it")
    ap.add_argument("--workload", default="replicas_sphere")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--mols-per-replica", type=int, default=150000)
    ap.add_argument("--subpart", type=float, default=None,
                    help="subpartition edge in um (default: the workload's, or the autotuner's for replicas)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-agreement", action="store_true",
                    help="skip the FP32-vs-FP64 hit-index comparison (wall path, 20 iterations)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device-to-device copy (known HBM bytes) to scale rocprof counters")
    return ap.parse_args()


def copy_calibration():
    """A plain copy of known size: 1 GiB read + 1 GiB written per launch."""
    import torch
    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    secs = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "launches": reps,
                      "GBps": 2 * (1 << 30) / secs / 1e9}), flush=True)


def build_workload(args):
    from mcell_amd import workloads as W
    if args.workload == "free_diffusion_replicas":
        wl = W.free_diffusion_replicas(args.replicas, args.mols_per_replica, iterations=args.steps)
        cfg = {"workload": "free diffusion, %d replicas x %d A in a shared reflective 1 um cube (12 walls), "
                           "dt 1e-6 s (subpartition edge: subpartition_um)" % (args.replicas, args.mols_per_replica),
               "replicas": args.replicas, "molecules": wl.num_molecules, "walls": int(len(wl.walls))}
    elif args.workload == "replicas_sphere":
        wl = W.replicas_sphere_in_cube(num_replicas=args.replicas, iterations=args.steps)
        cfg = {"workload": "%d replicas of A + B -> C (1e8 /M/s) over one mesh: reflective 2 um cube + icosphere "
                           "r 0.5 um (5,120 triangles); per replica N_A ~ U[1e3, 2e5], N_B = N_A/2, "
                           "D ~ U[1e-7, 5e-6] cm^2/s (seed 4 + i), C with D 5e-7, counts every 10 iterations, "
                           "dt 1e-6 s" % args.replicas,
               "replicas": args.replicas, "molecules": wl.num_molecules, "walls": int(len(wl.walls))}
    elif args.workload == "volume_surface":
        wl = W.volume_surface(iterations=args.steps)
        cfg = {"workload": "L' + R' -> LR' (1e8 /M/s) on a sphere: reflective 2 um cube + icosphere r 0.5 um "
                           "(5,120 triangles), 20,000 receptors R' uniform by area, 100,000 L (D 1e-6 cm^2/s) "
                           "outside the sphere, 10,000 T crossing it through a transparent surface class, "
                           "0.1 um subpartitions, dt 1e-6 s", "molecules": wl.num_molecules,
               "surface_molecules": int(len(wl.surface_mols["ids"])), "walls": int(len(wl.walls)), "replicas": 1}
    elif args.workload == "crowded_spheres":
        wl = W.crowded_spheres(iterations=args.steps)
        cfg = {"workload": "64 perturbed icospheres (20,480 triangles each, radii U[0.1, 0.3] um, +-2% radial "
                           "vertex noise) in a reflective 4 um cube, 10M volume molecules A:B:E = 5:4:1 (D 1e-6, "
                           "1e-6, 2e-7 cm^2/s), A + B -> E (1e8 /M/s), E -> A + B (1e3 /s), 200,000 S' on the "
                           "spheres, A + S' -> AS' (5e7 /M/s), 0.05 um subpartitions, dt 1e-6 s",
               "molecules": wl.num_molecules, "surface_molecules": int(len(wl.surface_mols["ids"])),
               "walls": int(len(wl.walls)), "replicas": 1}
    elif args.workload == "bimol_box":
        wl = W.bimol_box(iterations=args.steps)
        cfg = {"workload": "A + B -> C (1e8 /M/s) in a reflective 1 um cube (12 walls), 50,000 A + 50,000 B "
                           "(D 1e-6 cm^2/s), C with D 5e-7, 0.05 um subpartitions, counts every 10 iterations, "
                           "dt 1e-6 s", "molecules": wl.num_molecules, "walls": int(len(wl.walls)), "replicas": 1}
    elif args.workload == "free_diffusion_box":
        wl = W.free_diffusion_box(iterations=args.steps)
        cfg = {"workload": "free diffusion, 10,000 A in a reflective 1 um cube, 0.1 um subpartitions",
               "molecules": wl.num_molecules, "walls": int(len(wl.walls))}
    else:
        raise SystemExit("unknown workload %s" % args.workload)
    return wl, cfg


def cpu_baseline(seconds, workload):
    """The oracle in reference mode (ISAAC64 global stream, sequential FP64)
    for about `seconds` of CPU time: the reference's free-diffusion case, or
    for the reaction workloads the single-instance A + B -> C cube."""
    from mcell_amd import workloads as W
    from oracle import pyoracle as po
    if workload.startswith("free"):
        wl = W.free_diffusion_box()
        what = "free diffusion of 10,000 A in the reflective 1 um cube"
    elif workload == "volume_surface":
        wl = W.volume_surface()
        what = "the whole volume-surface case (110,000 volume molecules, 20,000 receptors, 5,132 walls)"
    elif workload == "crowded_spheres":
        wl = W.crowded_spheres(n_molecules=200_000)
        what = ("the crowded geometry (1,310,732 walls, 200,000 S') with 200,000 volume molecules A:B:E = 5:4:1 "
                "(1/50 of the GPU run's)")
    elif workload == "replicas_sphere":
        wl = W.replicas_sphere_in_cube(num_replicas=1)
        what = ("instance 0 of the replica batch (A + B -> C, %d molecules, sphere-in-cube mesh, 5,132 walls), "
                "one single-threaded process" % wl.num_molecules)
    else:
        wl = W.bimol_box()
        what = "A + B -> C, 50,000 A + 50,000 B in the reflective 1 um cube"
    w = po.world_from_workload(wl, draw_mode=po.DRAW_ISAAC)
    w.set_barrier_interval(wl.barrier_interval)
    done, secs, steps = 0, 0.0, 0
    chunk = 10
    while secs < seconds and done < 20000:
        secs += w.run(chunk)
        done += chunk
    steps = w.stats()["diffuse_3d_calls"]
    rate = steps / secs
    return {"value": rate, "unit": "molecule-timesteps/s", "cores": 1, "kind": "port",
            "sample": "oracle (FP64 restatement of the reference CPU path, ISAAC64 stream), %s, %d iterations, "
                      "%.1f s" % (what, done, secs)}


def fp32_hit_agreement(wl, iterations=20, max_molecules=500000):
    """FP32+fallback against FP64 on the workload's geometry and molecules
    without its reaction classes (the wall path: reflections and surface
    classes), per molecule-step: the walls hit in order (hit traces; the FP64
    form equals the oracle bit for bit), and the position RMS at the end.
    Batches use their first replica, large instances their first
    max_molecules molecules."""
    import dataclasses
    import numpy as np
    from mcell_amd.engine import DiffuseReactEngine
    sel = np.nonzero(wl.replica == 0)[0][:max_molecules]
    from mcell_amd import workloads as W
    sub = dataclasses.replace(wl, ids=wl.ids[sel], mol_species=wl.mol_species[sel], positions=wl.positions[sel],
                              replica=np.zeros(len(sel), np.uint32), num_replicas=1, replica_D=None,
                              species=wl.species_for(0), species_steps=W.species_steps(wl, 0))
    free = dataclasses.replace(sub, rxns=[], unimol_rxns=[], volsurf_rxns=[], surface_mols=None,
                               species_steps=[(ss, ts, f & ~5) for ss, ts, f in sub.species_steps])
    eng = {p: DiffuseReactEngine.from_workload(free, precision=p, trace=True) for p in ("fp64", "fp32")}
    steps = hits = differ = 0
    for it in range(iterations):
        tr = {}
        for p, e in eng.items():
            e.step(it, 1)
            m = e.molecules()
            w, c = e.hit_trace()
            o = np.argsort(m["id"], kind="stable")
            tr[p] = (w[o], c[o], m["pos"][o])
        (w64, c64, p64), (w32, c32, p32) = tr["fp64"], tr["fp32"]
        k = np.minimum(np.maximum(c64, c32), w64.shape[1])
        hit = np.maximum(c64, c32) > 0
        same = c64 == c32
        for j in range(w64.shape[1]):
            same &= (j >= k) | (w64[:, j] == w32[:, j])
        steps += len(c64)
        hits += int(hit.sum())
        differ += int((hit & ~same).sum())
    rms = float(np.sqrt(np.mean(np.sum((p64 - p32) ** 2, axis=1))))
    for e in eng.values():
        e.close()
    return {"iterations": iterations, "molecules": int(len(sel)), "molecule_steps": steps, "steps_with_hits": hits,
            "hit_index_disagreements": differ, "disagreement_rate": differ / max(hits, 1),
            "position_rms_lu": rms}


def reactive(wl):
    return bool(wl.rxns or wl.unimol_rxns or wl.volsurf_rxns)


def kernel_name(precision, wl):
    if reactive(wl):
        return "k_move_light + k_move_drain <%s>" % ("double" if precision == "fp64" else "float")
    return "k_light64 + k_hard64" if precision == "fp64" else "k_light32 + k_hard32"


def run_gpu(args, wl, precision):
    """Warm-up, then K timed iterations of one precision form; returns the
    step statistics and timings."""
    from mcell_amd.engine import DiffuseReactEngine
    eng = DiffuseReactEngine.from_workload(wl, precision=precision,
                                           capacity=wl.num_molecules * (2 if reactive(wl) else 1))
    print("bench: %s, %d molecules, precision %s, warmup %d, steps %d" % (args.workload, wl.num_molecules, precision,
                                                                          args.warmup, args.steps),
          file=sys.stderr, flush=True)
    if args.warmup:
        eng.step(0, args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    st = eng.step(args.warmup, args.steps)
    eng.sync()
    secs = time.perf_counter() - t0
    avg_ms, launches = eng.last_kernel_time()
    eng.close()
    steps_done = st.molecule_steps if reactive(wl) else wl.num_molecules * args.steps
    return {"secs": secs, "steps_done": steps_done, "value": steps_done / secs, "avg_ms": avg_ms,
            "reactions": st.reactions, "rounds": st.settle_rounds, "hard": st.slow_path_last,
            "reflections": st.wall_hits}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: the engine drives exactly one GPU (no sharding across GPUs); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    from mcell_amd import workloads as W
    wl, cfg = build_workload(args)
    main_prec = "fp32" if args.precision in ("fp32", "both") else "fp64"
    if args.subpart is not None:
        wl = W.with_subpartition(wl, args.subpart)
    elif wl.num_replicas > 1 and not args.no_autotune:
        from mcell_amd import autotune
        best, rates = autotune.choose_subpartition(
            wl, main_prec, log=lambda m: print(m, file=sys.stderr, flush=True))
        wl = W.with_subpartition(wl, best)
        cfg["autotune"] = {"subpartition_um": best, "rates": {str(k): v for k, v in rates.items()}}
    cfg["subpartition_um"] = wl.config.subpartition_dimension
    r = run_gpu(args, wl, main_prec)
    per_step = r["steps_done"] / args.steps
    achieved = B_ALG[main_prec] * per_step / (r["avg_ms"] * 1e-3) / 1e9
    out = {
        "metric": "molecule-timesteps/s on 1 MI355X (% HBM peak); FP32+fallback vs FP64 speedup",
        "value": r["value"],
        "unit": "molecule-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["secs"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kernel_name(main_prec, wl), "kernel_avg_ms": r["avg_ms"],
                     "bytes_per_molecule_step": B_ALG[main_prec]},
    }
    if reactive(wl):
        out["reactions"] = r["reactions"]
        out["settle_rounds"] = r["rounds"]
        out["roofline"]["full_step_fraction"] = r["hard"] / max(per_step, 1)
    else:
        out["roofline"]["slow_path_fraction"] = r["hard"] / max(wl.num_molecules, 1)
    if args.precision == "both":
        r64 = run_gpu(args, wl, "fp64")
        out["fp64"] = {"value": r64["value"], "ms_per_step": 1e3 * r64["secs"] / args.steps,
                       "kernel_avg_ms": r64["avg_ms"], "reactions": r64["reactions"]}
        out["fp32_over_fp64"] = r["value"] / r64["value"]
    if not args.no_agreement:
        try:
            out["fp32_vs_fp64"] = fp32_hit_agreement(wl)
        except Exception as e:  # a side measurement; the benchmark line stands without it
            out["fp32_vs_fp64"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
