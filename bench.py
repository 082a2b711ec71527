"""Benchmark of the DEM time step on one MI355X.

Workload (default): the large polydisperse granular collapse - 4,194,304 discs,
log-normal radii (median 0.35, clip [0.175, 0.525]; the stated radii scaled by
0.35 so the discs fit the 4096 x 2048 box), a packed column 2048 mean diameters wide
deposited in contact on the floor (every disc resting on two others), gravity,
spring-dashpot contacts kn_eff = 1e5, restitution 0.5, mu = 0.5, dt = 1e-4.  The
column is advanced --settle steps untimed (past its first neighbour-list rebuild and
re-sort), the block size and cell side are tuned on that state, then --warmup and
--steps follow: the timed window lies in the first few hundred of the run's 5,000
steps, while the pile still holds about three contacts per disc (it loses them as the
soft pile falls and compacts).  A "step" is one full time step of every disc (LOM +
Sim of the reference, here: the step kernel, plus the neighbour-list rebuild chain -
binning, cell scan, scatter, list build - on the steps that need it, and the storage
re-sort when due).

Prints one JSON line.  `value` is particle-steps/s of the timed region (inputs
resident in HBM, host clock around K steps bracketed by device syncs).
`roofline` is for the dominant kernel (the fused step kernel): algorithmic bytes
per launch = N * (60 + 16 * mean contacts per particle) over its mean duration
from HIP events on the ctx's stream; `traffic` is the kernel's measured HBM bytes per
fused launch (counter summary under profiles/, default workload only).  The window's
mean contacts per disc, rebuilds and re-sorts are in `config.window`; a window with
less than one contact per disc is flagged there.  `cpu_baseline` times the CPU
oracle's restatement of the reference algorithm (AABB bucket grid, 14-slot table) on
a bounded sample of the same scene.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "physics-sim_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per fused step launch measured with the FETCH_SIZE / WRITE_SIZE counters
# (scaled by a copy of known size) on the default workload
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_step_plain_column_collapse_4M_traffic.md")


def measured_traffic(args, cell_size: float):
    """Bytes per launch of the step kernel from the committed counter summary, for a
    configuration that summary covers (4,194,304 discs, its settle, and one of its
    rows of block size and cell side); None otherwise."""
    if not (args.particles == 4194304 and args.per_row == 2048 and args.collision_interval == 1):
        return None
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    settle, rows = None, {}
    for line in open(TRAFFIC_PROFILE):
        w = line.split()
        if len(w) == 2 and w[0] == "settle:":
            settle = int(w[1])
        # "traffic block B cell_size C: bytes"
        if len(w) == 6 and w[0] == "traffic" and w[1] == "block" and w[3] == "cell_size":
            rows[(int(w[2]), float(w[4].rstrip(":")))] = float(w[5])
    if settle != args.settle:
        return None
    for (b, c), v in rows.items():
        if b == args.block and abs(c - cell_size) <= 1e-6:
            return v
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--settle", type=int, default=100,
                    help="untimed steps of the collapse before tuning, warm-up and timing")
    ap.add_argument("--particles", type=int, default=4194304)
    ap.add_argument("--per-row", type=int, default=2048)
    ap.add_argument("--cell", type=float, default=0.0, help="grid cell side (0: just above 2 r_max)")
    ap.add_argument("--reorder", type=int, default=256,
                    help="storage re-sort every k steps (with a neighbour-list rebuild)")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--collision-interval", type=int, default=1,
                    help="narrow phase every k-th step (the reference's collision_interval setting)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-particles", type=int, default=262144)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --block and --cell as given instead of tuning them on the workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (continuous batching) leg")
    ap.add_argument("--serve-scenes", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson arrival rate, scenes/s")
    ap.add_argument("--serve-substeps", type=int, default=50)
    ap.add_argument("--no-bonded", action="store_true", help="skip the 1M-disc bonded-assembly leg")
    ap.add_argument("--bonded-window", type=int, default=1000,
                    help="steps of the bonded leg's HIP-event window, also its sampling interval")
    ap.add_argument("--bonded-steps", type=int, default=50000,
                    help="steps of the bonded-assembly pull (50,000: strain 0.05)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also time a plain device copy of this many MiB (the reference kernel for the HBM "
                         "traffic counters and the achievable copy bandwidth)")
    return ap.parse_args()


def copy_calibration(mib, reps=5):
    """A plain device-to-device copy of `mib` MiB, `reps` times, through the HIP
    runtime the DEM library itself links (its blit copy kernel): known bytes read
    and written per launch, so that the FETCH_SIZE / WRITE_SIZE counters of a
    profiled run can be scaled to bytes.  Returns the mean launch time and the copy
    bandwidth (read + write bytes per second)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    nbytes = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError("%s failed: hip error %d" % (what, rc))
    ok(hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)), "hipMalloc")
    ok(hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)), "hipMalloc")
    try:
        ok(hip.hipMemset(src, 0x5a, C.c_size_t(nbytes)), "hipMemset")
        ok(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
        ok(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
        D2D = 3
        ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), D2D, None), "hipMemcpyAsync")
        ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        ok(hip.hipEventRecord(e0, None), "hipEventRecord")
        for _ in range(reps):
            ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), D2D, None), "hipMemcpyAsync")
        ok(hip.hipEventRecord(e1, None), "hipEventRecord")
        ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        probe = (C.c_ubyte * 64)()
        ok(hip.hipMemcpy(probe, C.c_void_p(dst.value + nbytes - 64), C.c_size_t(64), 2), "hipMemcpy")
        assert all(b == 0x5a for b in probe)
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    t = ms.value / reps
    return {"mib": mib, "bytes_read": nbytes, "bytes_written": nbytes, "ms": t, "reps": reps,
            "gbs_read_plus_write": 2 * nbytes / (t * 1e-3) / 1e9}


def serving_bench(args):
    """Continuous batching against plain one-scene-at-a-time serving on the same
    stream, drawn from the serving configuration's distribution (scenes.serving_stream:
    half granular boxes, half bonded chains struck by an r = 8 disc; n log-uniform in
    [1k, 256k]; 2,000-20,000 steps; seeded per-scene friction, restitution and bond
    strengths), --serve-scenes of them (the configuration's stream holds 2,000).  Each
    mode runs twice: with Poisson arrivals at --serve-rate (latency under that load)
    and with every scene queued at t = 0 (service capacity).  "plain" is the same
    server admitting one scene at a time, so the difference is the batching alone;
    both replay each tick's launches from HIP graphs."""
    from psim import scenes
    from psim.autotune import Tuner
    from psim.serve import Server
    stream = scenes.serving_stream(n_scenes=args.serve_scenes, rate=args.serve_rate)
    # cell sizes per scene class (kind, size bucket) at the server's block size, from the
    # default cell up: in a batch a list rebuild covers every resident scene, so the thin
    # skin of a smaller cell, which a scene alone may afford, makes the whole batch
    # rebuild more often
    tuner = Tuner(blocks=(256,), factors=(1.0, 1.25), steps=60, warmup=5)   # long enough to include list rebuilds
    grids = [tuner.grid_for("box" if i % 2 == 0 else "chain", s, st, block=256)
             for i, (_, s, st, _) in enumerate(stream)]
    slots = sum((st.p_count + 255) // 256 * 256 for _, _, st, _ in stream)
    bonds = sum(st.n_bonds for _, _, st, _ in stream)
    ps_total = sum(st.p_count * n for _, _, st, n in stream)

    import math
    cells = sum(g.w * g.h + 1 for g in grids)

    def nearest_rank(sorted_vals, f):
        k = min(len(sorted_vals), max(1, math.ceil(f * len(sorted_vals))))
        return sorted_vals[k - 1]

    def run(max_resident, saturated):
        srv = Server(slot_capacity=slots, cell_capacity=cells, bond_capacity=max(bonds, 1), scene_capacity=len(stream),
                     substeps=args.serve_substeps, max_resident=max_resident, record_data=False, graph=True)
        reqs = [(0.0 if saturated else t, s, st.copy(), n, g) for (t, s, st, n), g in zip(stream, grids)]
        _, lat, wall = srv.run_stream(reqs)
        stt = srv.stats()
        srv.close()
        lat = sorted(lat)
        return {"scenes_per_s": len(reqs) / wall, "p50_s": nearest_rank(lat, 0.50), "p99_s": nearest_rank(lat, 0.99),
                "wall_s": wall, "failed": int(stt["failed"]),
                "particle_steps_per_s": ps_total / wall, "ticks": int(stt["ticks"])}

    out = {"stream": {"scenes": len(stream), "arrival_rate": args.serve_rate, "particles_total":
                      sum(st.p_count for _, _, st, _ in stream), "particle_steps_total": ps_total,
                      "substeps_per_tick": args.serve_substeps,
                      "tuned_cell_factors": {"%s/2^%d" % k: v["cell_factor"] for k, v in tuner.cache.items()},
                      "graphs": True,
                      "mix": "50/50 granular boxes (f32 contacts) / bonded chains of 256 discs struck by an r = 8 disc "
                             "(parallel bonds, Weibull(6) strengths); n log-uniform [1024, 262144], steps U[2000, 20000], "
                             "mu U[0.2, 0.8], restitution U[0.2, 0.9], strength scale U[20, 80]; %d scenes of the "
                             "configuration's 2,000" % len(stream)}}
    for name, mr in (("batched", 0), ("plain", 1)):
        out[name] = {"poisson": run(mr, False), "saturated": run(mr, True)}
        print("bench: serving %s done" % name, file=sys.stderr, flush=True)
    out["speedup_scenes_per_s_saturated"] = (out["batched"]["saturated"]["scenes_per_s"] /
                                             out["plain"]["saturated"]["scenes_per_s"])
    out["p99_ratio_poisson_plain_over_batched"] = out["plain"]["poisson"]["p99_s"] / out["batched"]["poisson"]["p99_s"]
    return out


def bonded_bench(args):
    """The bonded-assembly tensile test at 1,048,576 discs (scenes.bonded_assembly: a
    seeded triangular lattice of r ~ U[0.8, 1.2] discs at area fraction 0.89 relaxed
    on the device, every pair closer than 0.08 = 0.1 r_min bonded by the device bond
    builder, parallel bonds with Weibull(4, 30) strengths, bottom and top rows gripped
    and pulled apart at 1e-3 strain per 1,000 steps): --bonded-steps steps (50,000:
    strain 0.05) with the stress-strain trace and the broken-bond count sampled from
    the device every --bonded-window steps (1,000: dem_group_sums on the two grips,
    dem_broken_bond_count), the first window under HIP-event timing for the per-kernel
    times and the roofline of the bond pass and the bonded step kernel."""
    from psim import scenes
    from psim.compute import ComputeProg
    t = time.perf_counter()
    s, st, bottom, top, info = scenes.bonded_assembly(
        progress=lambda d, tot: print("bench: bonded relaxation %d / %d steps" % (d, tot), file=sys.stderr, flush=True))
    t_build = time.perf_counter() - t
    n, nb = st.p_count, info["bonds"]
    prog = ComputeProg(s, st, bond_capacity=nb, record_data=False, reorder_interval=args.reorder)
    every = max(1, args.bonded_window)
    trace = []

    def sample(done):
        gb, gt = prog.group_sums(bottom), prog.group_sums(top)
        strain = (gt["pos"][1] / gt["count"] - gb["pos"][1] / gb["count"]) / info["L0"] - 1.0
        stress = 0.5 * (gb["force"][1] - gt["force"][1]) / info["width"]
        trace.append([done, strain, stress, prog.broken_bond_count()])

    # the first window under HIP-event timing (the bonds still intact: the window
    # of the per-kernel figures), the rest timed by the host clock
    c_a = prog.counters()
    prog.timing(True)
    prog.step(every)
    prog.sync()
    tm = prog.timing(False)
    c_b = prog.counters()
    print("bench: bonded event window done (%d steps)" % every, file=sys.stderr, flush=True)
    sample(every)
    done = every
    prog.sync()
    t0 = time.perf_counter()
    while done < args.bonded_steps:
        k = min(every, args.bonded_steps - done)
        prog.step(k)
        done += k
        sample(done)
        if args.bonded_steps <= 10000 or done % 10000 == 0:
            print("bench: bonded pull %d / %d steps" % (done, args.bonded_steps), file=sys.stderr, flush=True)
    prog.sync()
    wall = time.perf_counter() - t0
    c0 = prog.counters()
    broken = prog.broken_bond_count()
    prog.close()
    timed_steps = max(args.bonded_steps - every, 1)
    nc = c_b["contacts"] / n
    nrec = c_b["bond_endpoints"]
    k_ms = tm["step_ms"] / max(tm["n_step"], 1)
    b_ms = tm["bonds_ms"] / max(tm["n_bonds"], 1)
    # algorithmic bytes: the bond pass reads and writes each live record's persistent
    # state {ft, fbt, theta_b} (2 x 12 B) and reads its bond's rest length and strengths
    # (8 B, shared by the endpoints); the step kernel reads and writes each disc's state
    # (48 B read, 32 B written incl. the contact count) and its contacts (16 B each).
    rec_bytes = 32.0 * nrec
    step_bytes = n * (60.0 + 16.0 * nc)
    peak = 8.0e12
    peak_unit = 1e9
    return {"workload": "bonded-assembly tensile test, %d discs r ~ U[0.8, 1.2] (lattice at area fraction 0.89 "
                        "relaxed on the device), %d parallel bonds (gap < 0.08, Weibull(4, 30) strengths, "
                        "kn = ks = 1500), grips pulled at 1e-3 strain per 1,000 steps, dt 1e-4, f32"
                        % (n, nb),
            "value": n * timed_steps / wall, "unit": "particle-steps/s", "steps": args.bonded_steps,
            "ms_per_step": 1e3 * wall / timed_steps,
            "timed": "steps 1,000 to the end by the host clock, with the trace sampling",
            "scene_build_s": t_build,
            "bonds": nb, "bonds_per_disc": 2.0 * nb / n, "grip_separation": info["L0"],
            "specimen_width": info["width"], "v_grip": info["v_grip"],
            "broken_bonds_end": int(broken), "list_rebuilds": c0["broad_phases"], "resorts": c0["reorders"],
            "window": {"steps": every, "first_step": 0, "live_bond_endpoints": nrec,
                       "mean_contacts_per_particle": nc,
                       "list_rebuilds": c_b["broad_phases"] - c_a["broad_phases"],
                       "kernel_ms_per_step": {"bond_pass": b_ms, "step": k_ms,
                                              "rest": (tm["total_ms"] - tm["step_ms"] - tm["bonds_ms"]) / every}},
            "roofline": {"bond_pass": {"bound": "hbm", "achieved": rec_bytes / (b_ms * 1e-3) / peak_unit,
                                       "peak": peak / peak_unit, "unit": "GB/s",
                                       "frac": rec_bytes / (b_ms * 1e-3) / peak,
                                       "alg_bytes": rec_bytes},
                         "step": {"bound": "hbm", "achieved": step_bytes / (k_ms * 1e-3) / peak_unit,
                                  "peak": peak / peak_unit, "unit": "GB/s", "frac": step_bytes / (k_ms * 1e-3) / peak,
                                  "alg_bytes": step_bytes}},
            "stress_strain": {"columns": ["step", "strain", "grip_stress", "broken_bonds"], "every": every,
                              "rows": trace}}


def graph_bench(args):
    """Repeated dem_step(50) calls on one small scene (a 16,384-disc granular box, the
    reference's frame loop: gen_per_frame steps per frame), launched directly and
    replayed from HIP graphs; microseconds per step over 100 calls."""
    from psim import scenes
    from psim.compute import ComputeProg
    out = {"scene": "granular box, 16384 discs, dem_step(50) x 100, re-sort off"}
    s, st = scenes.granular_box(n_side=128, seed=3, spacing=1.9)
    for g in (False, True):
        prog = ComputeProg(s, st.copy(), reorder_interval=-1, graph=g, record_data=False)
        for _ in range(4):
            prog.step(50)
        prog.sync()
        t = time.perf_counter()
        for _ in range(100):
            prog.step(50)
        prog.sync()
        dt = time.perf_counter() - t
        c = prog.counters()
        prog.close()
        out["graph" if g else "direct"] = {"us_per_step": dt / 5000 * 1e6, "captures": c["graph_captures"],
                                           "replays": c["graph_replays"]}
    out["speedup"] = out["direct"]["us_per_step"] / out["graph"]["us_per_step"]
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args):
    """Oracle (reference algorithm, f32) on a bounded sample: a `cpu_particles`-disc
    column of the same generator and width (the bottom layers of the full pile, in
    contact the same way), timed over enough steps for about cpu_seconds of work on up
    to 16 host cores, and on one core for about a third of that."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C
    import numpy as np
    import orc
    from psim import scenes
    from psim.compute import ComputeProg
    n = args.cpu_particles
    s, st = scenes.column_collapse(n=n, per_row=args.per_row)
    grid = ComputeProg.default_grid(s, st)
    # reference grid semantics (AABB buckets with capacity) on the box-sized grid
    rmin, rmax = float(st.radii[:n].min()), float(st.radii[:n].max())
    grid.cell_cap = min(int(np.ceil((2 * rmax / rmin + 1.0) ** 2)), n) + 2
    u = s.collision_settings()
    u.particles = n
    L = orc.lib()

    def timed(cores, seconds):
        eng = orc.OracleEngine(st, s, mode=orc.MODE_REFERENCE, grid=grid)
        v = eng._view()
        args_ = (C.byref(u), eng.mats, len(eng.mats), C.byref(grid), C.byref(v))
        # calibrate with 2 steps, then time a bounded number of steps from there
        t = L.orc_time_reference(*args_, 2, cores, s.physics.restitution)
        steps = max(2, int(seconds / max(t / 2, 1e-6)))
        t = L.orc_time_reference(*args_, steps, cores, s.physics.restitution)
        return steps, t
    cores = min(16, os.cpu_count() or 1)
    steps, t = timed(cores, args.cpu_seconds)
    steps1, t1 = timed(1, args.cpu_seconds / 3.0)
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "value_1core": n * steps1 / t1, "cpu_model": cpu_model(),
            "sample": "CPU restatement of the reference algorithm (oracle, reference slot rules, f32, OpenMP, "
                      "contact dashpot e %.1f) on a %d-disc column of the same width, %d steps in %.1f s on %d "
                      "cores (%d steps in %.1f s on 1 core)" % (s.physics.restitution, n, steps, t, cores, steps1,
                                                                 t1)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (multi-GPU sharding is a "
              "design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from psim import scenes
    from psim.compute import ComputeProg

    s, st = scenes.column_collapse(n=args.particles, per_row=args.per_row)
    s.physics.collision_interval = args.collision_interval
    n = st.p_count
    if args.settle:
        # the collapse's first steps, untimed; results do not depend on the grid or block
        pre = ComputeProg(s, st, record_data=False, reorder_interval=args.reorder)
        pre.step(args.settle)
        st = pre.update_state().copy()
        c = pre.counters()
        pre.close()
        print("bench: settled %d steps, %.3f contacts/disc" % (args.settle, c["contacts"] / n), file=sys.stderr,
              flush=True)
    tuned = None
    if not args.no_autotune:
        from psim.autotune import cell_size_for, tune
        tuned = tune(s, st)
        args.block = tuned["block"]
        args.cell = cell_size_for(s, st, tuned["cell_factor"])
        print("bench: autotuned block %d, cell factor %.2f" % (args.block, tuned["cell_factor"]), file=sys.stderr,
              flush=True)
    grid = ComputeProg.default_grid(s, st, cell_size=args.cell if args.cell > 0 else None)
    t_up = time.perf_counter()
    prog = ComputeProg(s, st, grid=grid, record_data=False, reorder_interval=args.reorder, block_size=args.block)
    prog.sync()
    t_up = time.perf_counter() - t_up   # pool allocation + host layout conversion + PCIe upload
    print("bench: %d discs, grid %dx%d cell %.4f; warmup %d" % (n, grid.w, grid.h, grid.cell_size, args.warmup),
          file=sys.stderr, flush=True)
    if args.warmup:
        prog.step(args.warmup)
    prog.sync()
    c0 = prog.counters()
    # timed region: K steps, host clock, device synced on both sides
    prog.timing(False)
    prog.sync()
    t0 = time.perf_counter()
    prog.step(args.steps)
    prog.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    c1 = prog.counters()
    # same K steps again with HIP events around every launch (per-kernel means)
    prog.timing(True)
    prog.step(args.steps)
    prog.sync()
    tm = prog.timing(False)
    c2 = prog.counters()
    t_down = time.perf_counter()
    prog.update_state()                 # PCIe download + rebuild of the reference layouts
    t_down = time.perf_counter() - t_down
    nc = (c0["contacts"] + c1["contacts"] + c2["contacts"]) / 3.0 / n   # both passes' ends and start
    ms_step = 1e3 * wall / args.steps
    value = n * args.steps / wall
    cells = grid.w * grid.h
    rebuilds = c1["broad_phases"] - c0["broad_phases"]
    resorts = c1["reorders"] - c0["reorders"]
    # bytes per particle-step, whole step: the step kernel's 60 + 16 n_c every step, the
    # binning's 16 + 8 C/N on the steps whose lists were rebuilt
    b_alg = 48 + 12 + 16 * nc + (16 + 8.0 * cells / n) * rebuilds / args.steps
    k_bytes = n * (60 + 16 * nc)                            # dominant kernel, per launch
    k_ms = tm["step_ms"] / max(tm["n_step"], 1)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    if nc < 1.0:
        print("bench: WARNING: the timed window holds %.3f contacts per disc (< 1)" % nc, file=sys.stderr, flush=True)
    out = {
        "metric": "particle-steps/s at 4M discs (%HBM peak); scenes/s + p99 latency batched vs plain",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "granular column collapse, %d discs (log-normal r median 0.35 clip [0.175,0.525]), "
                        "packed column deposited in contact, 4096x2048 box, spring-dashpot contacts (e 0.5), dt 1e-4, "
                        "steps %d-%d of the run" % (n, args.settle + args.warmup, args.settle + args.warmup + args.steps),
            "grid": [grid.w, grid.h], "cell_size": grid.cell_size, "reorder_interval": args.reorder,
            "collision_interval": args.collision_interval,
            "block": args.block, "autotune": tuned, "mean_contacts_per_particle": nc,
            "window": {"first_step": args.settle + args.warmup, "steps": args.steps,
                       "mean_contacts_per_particle": nc, "rebuilds": rebuilds, "resorts": resorts,
                       "contact_rich": nc >= 1.0},
            "hbm_frac_whole_step": value * b_alg / (HBM_PEAK_GBS * 1e9),
            "bytes_per_particle_step_alg": b_alg,
            "upload_ms": 1e3 * t_up, "download_ms": 1e3 * t_down,
            "kernel_ms": {"step": k_ms,
                          "rebuild_chain_per_step": (tm["scan_ms"] + tm["scatter_ms"]) / args.steps,
                          "reorder_per_step": tm["reorder_ms"] / args.steps},
            "counters": {k: c2[k] - c0.get(k, 0)
                         if k in ("steps", "broad_phases", "reorders", "list_overflow") else c2[k]
                         for k in c2},
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args, grid.cell_size)},
    }
    prog.close()
    if not args.no_bonded:
        out["bonded"] = bonded_bench(args)
        print("bench: bonded assembly done", file=sys.stderr, flush=True)
    if not args.no_serve:
        out["serving"] = serving_bench(args)
        out["graphs"] = graph_bench(args)
    if args.copy_mib > 0:
        out["config"]["copy_calibration"] = copy_calibration(args.copy_mib)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
