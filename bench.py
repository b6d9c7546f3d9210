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
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) that fill roofline.traffic")
    ap.add_argument("--pmc-inner", action="store_true",
                    help="(run by the traffic passes under rocprofv3) a few iterations, then the 1 GiB copy")
    ap.add_argument("--cpu-processes", type=int, default=None,
                    help="concurrent single-threaded oracle processes for the replica batch's CPU baseline "
                         "(default: min(usable cores, 16))")
    return ap.parse_args()


COPY_BYTES = 1 << 30


def _hip():
    import ctypes
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    raise OSError("libamdhip64.so not found")


def copy_peak(reps=10, quiet=False):
    """A plain device copy of known size (1 GiB read + 1 GiB written per
    launch, hipMemcpy device to device): the measured HBM copy rate, and the
    reference launch for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = _hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(COPY_BYTES)
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("hipMalloc of the copy buffers failed")
    try:
        hip.hipMemset(a, 1, n)
        hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            hip.hipMemcpy(b, a, n, 3)
        hip.hipDeviceSynchronize()
        secs = (time.perf_counter() - t0) / reps
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    gbps = 2 * COPY_BYTES / secs / 1e9
    if not quiet:
        print(json.dumps({"copy_bytes_read": COPY_BYTES, "copy_bytes_written": COPY_BYTES, "launches": reps,
                          "GBps": gbps}), flush=True)
    return gbps


def copy_calibration():
    copy_peak()


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
                           "D ~ U[1e-7, 5e-6] cm^2/s (seed 4 + i), C with D 5e-7, species counts of every "
                           "replica every 10 iterations, dt 1e-6 s" % args.replicas,
               "replicas": args.replicas, "molecules": wl.num_molecules, "walls": int(len(wl.walls))}
    elif args.workload == "volume_surface":
        wl = W.volume_surface(iterations=args.steps)
        cfg = {"workload": "L' + R' -> LR' (1e8 /M/s) on a sphere: reflective 2 um cube + icosphere r 0.5 um "
                           "(5,120 triangles), 20,000 receptors R' uniform by area, 100,000 L (D 1e-6 cm^2/s) "
                           "outside the sphere, 10,000 T crossing it through a transparent surface class, "
                           "0.1 um subpartitions, dt 1e-6 s", "molecules": wl.num_molecules,
               "surface_molecules": int(len(wl.surface_mols["ids"])), "walls": int(len(wl.walls)), "replicas": 1}
    elif args.workload == "volume_surface_diffusing":
        wl = W.with_diffusing_surface(W.volume_surface(iterations=args.steps), {1: 1e-8, 2: 1e-8})
        cfg = {"workload": "the volume-surface case with diffusing receptors: L' + R' -> LR' (1e8 /M/s), R' and LR' "
                           "diffusing on the icosphere r 0.5 um (5,120 triangles, D 1e-8 cm^2/s), 20,000 receptors, "
                           "100,000 L (D 1e-6 cm^2/s) outside the sphere, 10,000 T crossing it through a "
                           "transparent surface class, 0.1 um subpartitions, dt 1e-6 s",
               "molecules": wl.num_molecules, "surface_molecules": int(len(wl.surface_mols["ids"])),
               "walls": int(len(wl.walls)), "replicas": 1}
    elif args.workload == "volume_surface_diffusing_timesteps":
        wl = W.with_time_steps(W.with_diffusing_surface(W.volume_surface(iterations=args.steps), {1: 1e-8, 2: 1e-8}),
                               [1, 0.5, 0.5, 1])
        cfg = {"workload": "the volume-surface case with diffusing receptors taking time steps of 1/2 iteration: "
                           "L' + R' -> LR' (1e8 /M/s), R' and LR' diffusing on the icosphere r 0.5 um (5,120 "
                           "triangles, D 1e-8 cm^2/s, two steps per iteration), 20,000 receptors, 100,000 L "
                           "(D 1e-6 cm^2/s) outside the sphere, 10,000 T crossing it through a transparent surface "
                           "class, 0.1 um subpartitions, dt 1e-6 s",
               "molecules": wl.num_molecules, "surface_molecules": int(len(wl.surface_mols["ids"])),
               "walls": int(len(wl.walls)), "replicas": 1}
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
    elif args.workload == "bimol_box_timesteps":
        wl = W.with_time_steps(W.bimol_box(iterations=args.steps), [0.5, 1, 3])
        cfg = {"workload": "A + B -> C (1e8 /M/s) in a reflective 1 um cube (12 walls), 50,000 A + 50,000 B "
                           "(D 1e-6 cm^2/s) with species time steps of 1/2 (A, two steps per iteration), 1 (B) "
                           "and 3 iterations (C, D 5e-7, maturing from its reaction time), 0.05 um "
                           "subpartitions, counts every 10 iterations, dt 1e-6 s",
               "molecules": wl.num_molecules, "walls": int(len(wl.walls)), "replicas": 1}
    elif args.workload == "free_diffusion_box":
        wl = W.free_diffusion_box(iterations=args.steps)
        cfg = {"workload": "free diffusion, 10,000 A in a reflective 1 um cube, 0.1 um subpartitions",
               "molecules": wl.num_molecules, "walls": int(len(wl.walls))}
    else:
        raise SystemExit("unknown workload %s" % args.workload)
    return wl, cfg


def host_cpu():
    """nproc, the cores this process may use, and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "usable_cores": usable, "model": model}


def _oracle_instance(job):
    """One single-threaded oracle process on one instance of the replica
    batch for about `seconds`: (molecule-steps, seconds)."""
    r, seconds, subpart = job
    sys.path.insert(0, ROOT)
    from mcell_amd import workloads as W
    from oracle import pyoracle as po
    wl = W.replicas_sphere_in_cube(num_replicas=r + 1)
    if subpart is not None:
        wl = W.with_subpartition(wl, subpart)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_ISAAC, replica=r)
    w.set_barrier_interval(wl.barrier_interval)
    secs = 0.0
    done = 0
    while secs < seconds and done < 20000:
        secs += w.run(10)
        done += 10
    steps = w.stats()["diffuse_3d_calls"]
    w.close()
    return steps, secs


def cpu_baseline_replicas(seconds, processes, subpart):
    """The replica batch on the CPU as MCell4 runs a sweep: one
    single-threaded process per instance, `processes` at a time
    (mcell4_runner.py -j): instances 0 .. processes-1, each for about
    `seconds`; the aggregate molecule-timesteps/s of the wave."""
    from multiprocessing import get_context
    jobs = [(r, seconds, subpart) for r in range(processes)]
    t0 = time.perf_counter()
    with get_context("spawn").Pool(processes) as pool:
        res = pool.map(_oracle_instance, jobs)
    wall = time.perf_counter() - t0
    steps = sum(r[0] for r in res)
    loop_secs = max(r[1] for r in res)
    one = res[0][0] / res[0][1]
    host = host_cpu()
    return {"value": steps / loop_secs, "unit": "molecule-timesteps/s", "cores": processes, "kind": "port",
            "processes": processes, "single_process_value": one, "host": host,
            "sample": "oracle (FP64 restatement of the reference CPU path, ISAAC64 stream), instances 0-%d of the "
                      "replica batch (A + B -> C, sphere-in-cube mesh, 5,132 walls), %d concurrent single-threaded "
                      "processes (one per instance, as mcell4_runner -j runs a sweep) on %d of %d cores (%s), about "
                      "%.0f s of iterations each (%.1f s wall for the wave)"
                      % (processes - 1, processes, processes, host["nproc"], host["model"], seconds, wall)}


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
    elif workload == "volume_surface_diffusing":
        wl = W.with_diffusing_surface(W.volume_surface(), {1: 1e-8, 2: 1e-8})
        what = ("the whole volume-surface case with diffusing receptors (110,000 volume molecules, 20,000 "
                "receptors, 5,132 walls)")
    elif workload == "volume_surface_diffusing_timesteps":
        wl = W.with_time_steps(W.with_diffusing_surface(W.volume_surface(), {1: 1e-8, 2: 1e-8}), [1, 0.5, 0.5, 1])
        what = ("the whole volume-surface case with diffusing receptors at time steps of 1/2 iteration (110,000 "
                "volume molecules, 20,000 receptors, 5,132 walls)")
    elif workload == "crowded_spheres":
        wl = W.crowded_spheres(n_molecules=200_000)
        what = ("the crowded geometry (1,310,732 walls, 200,000 S') with 200,000 volume molecules A:B:E = 5:4:1 "
                "(1/50 of the GPU run's)")
    elif workload == "replicas_sphere":
        wl = W.replicas_sphere_in_cube(num_replicas=1)
        what = ("instance 0 of the replica batch (A + B -> C, %d molecules, sphere-in-cube mesh, 5,132 walls), "
                "one single-threaded process" % wl.num_molecules)
    elif workload == "bimol_box_timesteps":
        wl = W.with_time_steps(W.bimol_box(), [0.5, 1, 3])
        what = ("A + B -> C, 50,000 A + 50,000 B in the reflective 1 um cube, time steps 1/2, 1 and 3 "
                "iterations")
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
    host = host_cpu()
    return {"value": rate, "unit": "molecule-timesteps/s", "cores": 1, "kind": "port", "host": host,
            "sample": "oracle (FP64 restatement of the reference CPU path, ISAAC64 stream), %s, %d iterations, "
                      "%.1f s, one single-threaded process on a %s (%d cores)"
                      % (what, done, secs, host["model"], host["nproc"])}


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


def counts_every(wl):
    """Iterations between species counts (the count barriers of the
    reaction workloads), or None."""
    return wl.barrier_interval if reactive(wl) and wl.barrier_interval > 1 else None


def run_gpu(args, wl, precision):
    """Warm-up, then K timed iterations of one precision form, with the
    species counts of the reaction workloads taken every count barrier
    (on the device, allocation-free) inside the timed region; returns the
    step statistics and timings."""
    import numpy as np
    from mcell_amd.engine import DiffuseReactEngine
    eng = DiffuseReactEngine.from_workload(wl, precision=precision,
                                           capacity=wl.num_molecules * (2 if reactive(wl) else 1))
    print("bench: %s, %d molecules, precision %s, warmup %d, steps %d" % (args.workload, wl.num_molecules, precision,
                                                                          args.warmup, args.steps),
          file=sys.stderr, flush=True)
    every = counts_every(wl)
    counts = np.zeros((wl.num_replicas, len(wl.species_steps)), np.uint64)
    if args.warmup:
        eng.step(0, args.warmup)
    eng.sync()
    it, end = args.warmup, args.warmup + args.steps
    kernel_ms = 0.0
    n_counts = 0
    st_sum = {"molecule_steps": 0, "reactions": 0, "settle_rounds": 0, "wall_hits": 0}
    hard = 0
    form = 0
    t0 = time.perf_counter()
    b = max(int(wl.barrier_interval), 1)
    while it < end:
        chunk = min(end - it, b - it % b)  # one mcg_step call: up to the next barrier
        st = eng.step(it, chunk)
        avg_ms, _ = eng.last_kernel_time()  # per iteration of that call
        kernel_ms += avg_ms * chunk
        for k in st_sum:
            st_sum[k] += getattr(st, k)
        hard = st.slow_path_last
        form = st.drain_form
        it += chunk
        if every and it % every == 0:
            eng.count_species_noalloc(counts)
            n_counts += 1
    eng.sync()
    secs = time.perf_counter() - t0
    eng.close()
    steps_done = st_sum["molecule_steps"] if reactive(wl) else wl.num_molecules * args.steps
    return {"secs": secs, "steps_done": steps_done, "value": steps_done / secs, "avg_ms": kernel_ms / args.steps,
            "reactions": st_sum["reactions"], "rounds": st_sum["settle_rounds"], "hard": hard,
            "reflections": st_sum["wall_hits"], "counts": n_counts,
            "drain_form": {0: None, 1: "exact_disk in place", 2: "exact_disk deferred"}.get(form)}


def traffic_passes(args, wl_subpart, main_prec, iterations):
    """HBM bytes per iteration of the dominant kernels from two rocprofv3
    --pmc passes over a short run of the same workload (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by what the counter reads for
    the 1 GiB copy in the same process.  None (with the reason) when the
    profiler is not available or fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    names = ("k_move_light<float>", "k_move_drain<float>") if main_prec == "fp32" else \
        ("k_move_light<double>", "k_move_drain<double>")
    wl_name = args.workload
    if wl_name not in ("replicas_sphere", "bimol_box", "crowded_spheres", "volume_surface",
                       "volume_surface_diffusing", "volume_surface_diffusing_timesteps", "bimol_box_timesteps"):
        names = ("k_light32", "k_mid32", "k_hard32") if main_prec == "fp32" else ("k_light64", "k_mid64", "k_hard64")
    out = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
                   sys.executable, os.path.join(ROOT, "bench.py"), "--pmc-inner", "--workload", wl_name,
                   "--steps", str(iterations), "--warmup", "1", "--precision", main_prec,
                   "--replicas", str(args.replicas), "--mols-per-replica", str(args.mols_per_replica)]
            if wl_subpart is not None:
                cmd += ["--subpart", str(wl_subpart)]
            print("bench: traffic pass %s" % counter, file=sys.stderr, flush=True)
            try:
                r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, timeout=300)
            except subprocess.TimeoutExpired:
                return None, "rocprofv3 %s pass timed out" % counter
            files = []
            for root, _, fs in os.walk(d):
                files += [os.path.join(root, f) for f in fs if f.endswith("counter_collection.csv")]
            if r.returncode != 0 or not files:
                return None, "rocprofv3 %s pass failed (rc %d)" % (counter, r.returncode)
            kern, copy = 0.0, []
            for row in csv.DictReader(open(files[0])):
                if row.get("Counter_Name") != counter:
                    continue
                name = row["Kernel_Name"]
                v = float(row["Counter_Value"]) * 1024.0  # KiB per dispatch
                if any(k in name for k in names):
                    kern += v
                elif "copy" in name.lower() or "elementwise" in name.lower():
                    copy.append(v)
            copy = [c for c in copy if c > 0.25 * COPY_BYTES]
            if not copy:
                return None, "no copy dispatch in the %s pass" % counter
            factor = COPY_BYTES / (sum(copy) / len(copy))
            out[counter] = (kern / (iterations + 1), factor)  # the warm-up iteration is profiled too
    fetch = out["FETCH_SIZE"][0] * out["FETCH_SIZE"][1]
    write = out["WRITE_SIZE"][0] * out["WRITE_SIZE"][1]
    return {"bytes_per_iteration": fetch + write, "fetch": fetch, "write": write,
            "fetch_factor": out["FETCH_SIZE"][1], "write_factor": out["WRITE_SIZE"][1], "kernels": list(names),
            "iterations_profiled": iterations + 1}, None


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
    if args.pmc_inner:
        if args.subpart is not None:
            wl = W.with_subpartition(wl, args.subpart)
        args.no_autotune = True
        run_gpu(args, wl, main_prec)
        copy_peak(reps=2, quiet=True)
        return
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
        out["drain_form"] = r["drain_form"]
        out["species_counts_taken"] = r["counts"]
        out["roofline"]["full_step_fraction"] = r["hard"] / max(per_step, 1)
    else:
        out["roofline"]["slow_path_fraction"] = r["hard"] / max(wl.num_molecules, 1)
    if args.precision == "both":
        r64 = run_gpu(args, wl, "fp64")
        out["fp64"] = {"value": r64["value"], "ms_per_step": 1e3 * r64["secs"] / args.steps,
                       "kernel_avg_ms": r64["avg_ms"], "reactions": r64["reactions"]}
        out["fp32_over_fp64"] = r["value"] / r64["value"]
    # the measured copy rate beside the published peak
    try:
        out["roofline"]["copy_peak_GBps"] = copy_peak(quiet=True)
    except Exception as e:  # a side measurement
        out["roofline"]["copy_peak_GBps"] = None
        out["roofline"]["copy_peak_error"] = str(e)
    # wall records and partner positions are read through the caches, not
    # staged in LDS (LDS holds the Ziggurat tables and the full step's
    # subpartition sets): the LDS share of wall/candidate reads is zero
    out["roofline"]["lds_hit_rate"] = 0.0
    if not args.no_traffic:
        t, why = traffic_passes(args, cfg["subpartition_um"], main_prec, 3)
        if t is not None:
            out["roofline"]["traffic"] = t["bytes_per_iteration"] / 1e9
            out["roofline"]["traffic_unit"] = "GB per iteration of the dominant kernels"
            out["roofline"]["traffic_detail"] = t
            out["roofline"]["traffic_over_algorithmic"] = t["bytes_per_iteration"] / (B_ALG[main_prec] * per_step)
        else:
            out["roofline"]["traffic_error"] = why
    if not args.no_agreement:
        try:
            out["fp32_vs_fp64"] = fp32_hit_agreement(wl)
        except Exception as e:  # a side measurement; the benchmark line stands without it
            out["fp32_vs_fp64"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        if args.workload == "replicas_sphere":
            procs = args.cpu_processes or max(1, min(host_cpu()["usable_cores"], 16, args.replicas))
            out["cpu_baseline"] = cpu_baseline_replicas(args.cpu_seconds, procs, cfg["subpartition_um"])
        else:
            out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
