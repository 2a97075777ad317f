"""Benchmark: points/s of the triangulation path on one MI355X.

Workload: 1,000,000 points i.i.d. U[0,1)^3, f32, seed 1 (numpy PCG64), already
resident in HBM when the timed region starts; the result (tets + adjacency)
stays in HBM (wtet_delaunay_resident).  One step = one whole triangulation:
normalize + sort, insertion rounds, flips, exact audit, compaction, and the
host repair of any face the GPU phase left non-Delaunay.

Reported: the filtered path (f32 filter -> f64 filter -> exact) as ``value``
and the plain exact-only path beside it, the live roofline of the dominant
kernel (HIP events on the library's stream), and the CPU oracle (serial
Bowyer-Watson) timed on this host as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes of each kernel group over one call, from the run's own
# counters (minimal SoA traffic per unit):
#   flip_check (FP32-only check): per tet checked -- own quad + adjacency 32,
#     four neighbour quads 64, eight vertex xyz 96, vote atomic 8 = 200
#   flip_decide_exec: per tet checked 24 (its vote word, the face's vote
#     word, info) + per executed flip 208 (read/write <= 3 tets 192, flip
#     record 16)
#   flip_update: per executed flip 48 (six external back-pointers 24 and
#     their messages 24); decide_exec + update = 24 per check + 256 per flip
#   flip_small: the three above for the short-list iterations
#   split: per insertion 224; output: per tet 32
#   audit: per live tet 124 (quad 16, adjacency 32, info 4, its four points
#     48, the opposite points of the two faces it owns 24); per slot a 3-2
#     flip freed, its info word 4
def kernel_bytes(st):
    a_big = st["sum_active"] - st["small_checked"]
    f_big = st["sum_flips"] - st["small_flips"]
    inserted = st["num_points"] - st["points_left"] - st["duplicates"]
    return {
        "flip_check": 200 * a_big,
        "flip_decide_exec": 24 * a_big + 208 * f_big,
        "flip_update": 48 * f_big,
        "flip_small": 224 * st["small_checked"] + 256 * st["small_flips"],
        "split": 224 * inserted,
        "audit": 124 * st["num_tets_out"] + 4 * max(st["tet_slots"] - st["num_tets_out"], 0),
        "output": 32 * st["num_tets_out"],
    }


KERNEL_NAMES = {
    "flip_check": "k_flip_check<LEAN> (FP32-only flip check)",
    "flip_decide_exec": "k_flip_decide_exec (flip vote, claim and 2-3 / 3-2 execution)",
    "flip_update": "k_flip_update (external adjacency through flip messages)",
    "flip_small": "k_flip_small (one-workgroup iterations of short lists)",
    "split": "k_split (1-4 split)",
    "audit": "k_audit<LEAN> (exact audit of every finite face)",
    "output": "k_output (compaction and id remap)",
}


def job_bytes(st):
    """Whole-job algorithmic bytes from the run's own counters:
    12N (read xyz) + 128 per uninserted point per round + 224N (splits)
    + 200 per checked tet + 256 per flip + 32 per output tet."""
    n = st["num_points"]
    return (12 * n + 128 * st["sum_uninserted"] + 224 * n + 200 * st["sum_active"]
            + 256 * st["sum_flips"] + 32 * st["num_tets_out"])


def host_info():
    """Core count, CPU model and the oracle's compiler flags of this host."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    flags = None
    try:
        for line in open(os.path.join(ROOT, "oracle", "Makefile")):
            if line.startswith("CFLAGS"):
                flags = line.split("=", 1)[1].strip()
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "compiler": "gcc", "cflags": flags}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the exact-only comparison")
    ap.add_argument("--ring", type=int, default=1, help="insertion independence radius (0, 1, 2)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4,096-cloud batched leg")
    ap.add_argument("--no-stress", action="store_true", help="skip the lattice / sphere leg")
    ap.add_argument("--no-presize", action="store_true", help="skip the learned-presizing leg")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the 10,000-point Hilbert exact-only leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default entries per thread instead of autotuning them")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the flip iterations directly instead of replaying HIP graphs "
                         "(the same kernels; used under rocprofv3)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--config1-only", action="store_true",
                    help="only the 10,000-point Hilbert workload, exact only then filtered: 2 warmup "
                         "and 5 timed resident calls each, one JSON line per mode (the program "
                         "profiled for that workload)")
    ap.add_argument("--stress-only", default="", metavar="NAMES",
                    help="only the stress inputs named (comma-separated: lattice_128, sphere_1M): "
                         "one untimed and 2 timed resident calls each, one JSON line per input "
                         "(the program profiled for those workloads)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the streaming copy kernel over MIB MiB (5 launches), "
                         "the known-size reference for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the path is single-device and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(1)

    import wtet_amd
    from wtet_amd import workloads as W

    if args.copy_calibration:
        ctx = wtet_amd.Context(0)
        nbytes = args.copy_calibration << 20
        src = ctx.upload(np.zeros(nbytes, np.uint8))
        dst = ctx.upload(np.zeros(nbytes, np.uint8))
        t0 = time.perf_counter()
        for _ in range(5):
            ctx.copy(dst, src, nbytes)
        dt = (time.perf_counter() - t0) / 5
        ctx.free(src)
        ctx.free(dst)
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": round(dt * 1e3, 3),
                          "GBs_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}), flush=True)
        return

    if args.stress_only:
        ctx = wtet_amd.Context(0)
        ctx.set_flip_graphs(not args.no_graphs)
        inputs = {"lattice_128": lambda: W.lattice(128), "sphere_1M": lambda: W.sphere(1_000_000, 2)}
        for name in args.stress_only.split(","):
            pts_s = inputs[name]()
            d_s = ctx.upload(pts_s)
            ctx.delaunay_resident(d_s, len(pts_s))
            t = []
            for _ in range(2):
                t0 = time.perf_counter()
                ctx.delaunay_resident(d_s, len(pts_s))
                t.append((time.perf_counter() - t0) * 1e3)
            s1 = ctx.stats()
            ctx.free(d_s)
            print(json.dumps({"workload": name, "points": len(pts_s), "ms": [round(x, 2) for x in t],
                              "ms_device": round(s1["ms_device"], 3), "ms_repair": round(s1["ms_repair"], 3),
                              "rounds": s1["rounds"], "flip_iterations": s1["flip_iterations"],
                              "flips": s1["sum_flips"], "checked_tets": s1["sum_active"],
                              "violations_gpu": s1["violations_gpu"]}), flush=True)
        return

    if args.config1_only:
        ctx = wtet_amd.Context(0)
        p1 = W.uniform(10_000, 0)
        d_1 = ctx.upload(p1)
        for mode in ("exact", "filtered"):
            cfg = wtet_amd.Config(predicate_mode=mode, enable_hilbert_sorting=True, collect_stats=True)
            for _ in range(2):
                ctx.delaunay_resident(d_1, len(p1), cfg)
            t = []
            for _ in range(5):
                t0 = time.perf_counter()
                ctx.delaunay_resident(d_1, len(p1), cfg)
                t.append((time.perf_counter() - t0) * 1e3)
            s1 = ctx.stats()
            print(json.dumps({"workload": "uniform_10000_seed0", "order": "hilbert", "predicates": mode,
                              "ms": [round(x, 2) for x in t], "ms_device": round(s1["ms_device"], 3),
                              "rounds": s1["rounds"], "flip_iterations": s1["flip_iterations"],
                              "checked_tets": s1["sum_active"], "violations_gpu": s1["violations_gpu"]}),
                  flush=True)
        ctx.free(d_1)
        return

    pts = W.uniform(args.n, args.seed)
    ctx = wtet_amd.Context(0)
    d_pts = ctx.upload(pts)  # resident in HBM before any timed region
    ctx.set_insertion_ring(args.ring)
    ctx.set_flip_graphs(not args.no_graphs)
    filt = wtet_amd.Config(predicate_mode="filtered")
    exact = wtet_amd.Config(predicate_mode="exact")

    # counters and parity run (untimed)
    res = ctx.delaunay(pts, wtet_amd.Config(predicate_mode="filtered", collect_stats=True))
    st = res.stats
    pred = st["pred"]
    o3 = max(1, sum(pred[0:4]))
    ins = max(1, sum(pred[4:8]))
    stages = {
        "orient3d": {"f32": pred[0] / o3, "f64": pred[1] / o3, "exact": pred[2] / o3, "sos": pred[3] / o3},
        "insphere": {"f32": pred[4] / ins, "f64": pred[5] / ins, "exact": pred[6] / ins, "sos": pred[7] / ins},
    }

    def run(cfg, k):
        # every call returns after its own stream synchronization
        t0 = time.perf_counter()
        for _ in range(k):
            ctx.delaunay_resident(d_pts, args.n, cfg)
        return (time.perf_counter() - t0) * 1e3 / max(k, 1)

    # autotuned persistent grid sizing (entries per thread), outside the timed region
    items = ctx.flip_items()[0] if args.no_autotune else ctx.autotune(d_pts, [0, args.n], filt)[0]
    run(filt, args.warmup)
    ms = run(filt, args.steps)
    value = args.n / (ms / 1e3)
    print(f"filtered: {ms:.2f} ms/step, {value:.4g} points/s", file=sys.stderr, flush=True)
    st_t = ctx.stats()

    # the same step with the flip iterations launched directly instead of
    # replayed from HIP graphs (the form profiled under rocprofv3)
    direct = None
    if not args.no_graphs:
        ctx.set_flip_graphs(False)
        run(filt, 1)
        wall, dev = [], []
        for _ in range(3):
            wall.append(run(filt, 1))
            dev.append(ctx.stats()["ms_device"])
        ctx.set_flip_graphs(True)
        direct = {"ms_per_step_median": round(float(np.median(wall)), 3),
                  "ms_device_median": round(float(np.median(dev)), 3)}

    # end to end: host points in, host tets and adjacency out (PCIe both ways)
    ctx.delaunay(pts, filt)
    t0 = time.perf_counter()
    e2e_k = max(1, min(args.steps, 3))
    for _ in range(e2e_k):
        r_e = ctx.delaunay(pts, filt)
    ms_e2e = (time.perf_counter() - t0) * 1e3 / e2e_k
    end_to_end = {"value": round(args.n / (ms_e2e / 1e3), 1), "ms_per_step": round(ms_e2e, 3),
                  "steps": e2e_k, "bytes_in": int(pts.nbytes),
                  "bytes_out": int(r_e.tets.nbytes + r_e.adjacency.nbytes)}
    print(f"end to end (host in, host out): {ms_e2e:.2f} ms/step", file=sys.stderr, flush=True)

    # live per-kernel timing (HIP events on the library's stream)
    ctx.set_profiling(True)
    ctx.delaunay_resident(d_pts, args.n, filt)
    kt = ctx.kernel_times()
    st_p = ctx.stats()
    ctx.set_profiling(False)
    # roofline of the kernel group with the most measured time
    kb = kernel_bytes(st_p)
    per_kernel = {}
    for name, (launches, tot_ms) in kt.items():
        if not launches or name not in kb:
            continue
        gbs = kb[name] / (tot_ms / 1e3) / 1e9 if tot_ms > 0 else 0.0
        per_kernel[name] = {"launches": launches, "ms": round(tot_ms, 3),
                            "bytes_per_launch": int(kb[name] / launches),
                            "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms"])
    launches, tot_ms = kt[dom]
    achieved = per_kernel[dom]["achieved_GBs"]
    # measured HBM bytes per unit (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    # copy-kernel calibrated) times this run's units per launch
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "flip_kernels_traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath)).get(dom)
        if tj:
            units = {"flip_check": st_p["sum_active"] - st_p["small_checked"],
                     "flip_decide_exec": st_p["sum_flips"] - st_p["small_flips"],
                     "flip_update": st_p["sum_flips"] - st_p["small_flips"]}.get(dom)
            if units is not None:
                traffic = int(tj["bytes_per_unit"] * units / max(launches, 1))
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": KERNEL_NAMES[dom], "launches": launches,
                "avg_launch_ms": round(tot_ms / max(launches, 1), 4),
                "bytes_per_launch_avg": int(kb[dom] / max(launches, 1))}
    flip_counts = {"sum_active": st_p["sum_active"], "sum_flips": st_p["sum_flips"],
                   "flips_23": st_p["flips_23"], "flips_32": st_p["flips_32"],
                   "small_checked": st_p["small_checked"], "small_flips": st_p["small_flips"],
                   "per_point": {"checks": round(st_p["sum_active"] / args.n, 2),
                                 "flips": round(st_p["sum_flips"] / args.n, 2)}}

    plain = None
    if not args.no_plain:
        run(exact, 1)
        ms_plain = run(exact, max(1, min(args.steps, 3)))
        plain = {"value": args.n / (ms_plain / 1e3), "ms_per_step": round(ms_plain, 3)}
        print(f"exact-only: {ms_plain:.2f} ms/step", file=sys.stderr, flush=True)

    batch = None
    bp = boff = None
    if not args.no_batch:
        # 4,096 independent clouds (log-normal sizes, median 2,000, Gaussian
        # mixtures): one batched pass vs single-instance calls per cloud
        bp, boff = W.mixed_batch(4096, 4)
        d_b = ctx.upload(bp)
        nb = len(bp)
        items_b, _ = ctx.autotune(d_b, boff, filt)
        ctx.delaunay_batch_resident(d_b, boff, filt)
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            ctx.delaunay_batch_resident(d_b, boff, filt)
        ms_b = (time.perf_counter() - t0) * 1e3 / args.batch_steps
        bst = ctx.stats()
        print(f"batched: {ms_b:.2f} ms/step, {nb / (ms_b / 1e3):.4g} points/s", file=sys.stderr, flush=True)

        def singles(cfg, clouds):
            t0 = time.perf_counter()
            pts_done = 0
            for c in clouds:
                a, b = int(boff[c]), int(boff[c + 1])
                ctx.delaunay_resident(d_b + 12 * a, b - a, cfg)
                pts_done += b - a
            return pts_done / (time.perf_counter() - t0)

        sample = list(range(256))
        singles(filt, sample[:8])
        single_filt = singles(filt, sample)
        single_plain = singles(exact, sample)
        ctx.free(d_b)
        batch = {"workload": "mixed_batch_4096_seed4", "clouds": 4096, "points": nb,
                 "value": round(nb / (ms_b / 1e3), 1), "unit": "points/s", "ms_per_step": round(ms_b, 3),
                 "steps": args.batch_steps,
                 "single_instance_filtered": round(single_filt, 1),
                 "single_instance_plain_exact": round(single_plain, 1),
                 "single_instance_sample": "clouds 0..255, one resident call each",
                 "speedup_vs_single_plain": round(nb / (ms_b / 1e3) / single_plain, 2),
                 "autotuned_items_per_thread": items_b,
                 "ms_device": round(bst["ms_device"], 3), "ms_repair": round(bst["ms_repair"], 3),
                 "violations_gpu": bst["violations_gpu"], "rounds": bst["rounds"],
                 "flip_iterations": bst["flip_iterations"]}

    stress = None
    if not args.no_stress:
        # degenerate inputs: 128^3 lattice (2,097,152 points, exact fallback
        # and SoS on every cell) and 1M points on the unit sphere
        stress = {}
        for name, pts_s in (("lattice_128", W.lattice(128)), ("sphere_1000000_seed2", W.sphere(1_000_000, 2))):
            d_s = ctx.upload(pts_s)
            ctx.delaunay_resident(d_s, len(pts_s), wtet_amd.Config(collect_stats=True))
            ss = ctx.stats()
            # one untimed call with the timed configuration (the stats call
            # above runs other graphs), then the median of three calls
            ctx.delaunay_resident(d_s, len(pts_s), filt)
            runs, parts = [], []
            for _ in range(3):
                t0 = time.perf_counter()
                ctx.delaunay_resident(d_s, len(pts_s), filt)
                runs.append((time.perf_counter() - t0) * 1e3)
                st_s = ctx.stats()
                parts.append((st_s["ms_device"], st_s["ms_repair"], st_s["repair_retries"]))
            ms_s = float(np.median(runs))
            ctx.free(d_s)
            o3s, ins = max(1, sum(ss["pred"][0:4])), max(1, sum(ss["pred"][4:8]))
            stress[name] = {"points": len(pts_s), "ms": round(ms_s, 3),
                            "value": round(len(pts_s) / (ms_s / 1e3), 1),
                            "orient3d_exact_or_sos": (ss["pred"][2] + ss["pred"][3]) / o3s,
                            "insphere_exact_or_sos": (ss["pred"][6] + ss["pred"][7]) / ins,
                            "sos_insphere": ss["pred"][7] / ins,
                            "violations_gpu": ss["violations_gpu"], "duplicates": ss["duplicates"],
                            "ms_calls": [round(x, 3) for x in runs],
                            "ms_device_calls": [round(x[0], 3) for x in parts],
                            "ms_repair_calls": [round(x[1], 3) for x in parts],
                            "repair_retries_calls": [x[2] for x in parts]}
            print(f"stress {name}: {ms_s:.1f} ms", file=sys.stderr, flush=True)

    presized = None
    if not args.no_presize:
        # 5,000,000 points: 90% from 32 anisotropic Gaussians + 10% uniform;
        # each mode in a fresh context so allocation and growth are counted
        pp = W.anisotropic(5_000_000, 5)
        presized = {"workload": "anisotropic_5000000_seed5", "points": len(pp)}
        from wtet_amd import presize as PZ
        for mode in ("learned", "fixed", "grow"):
            if mode == "learned" and not os.path.exists(PZ.MODEL_PATH):
                continue
            c2 = wtet_amd.Context(0)
            d_p = c2.upload(pp)
            cfg_p = wtet_amd.Config(presize=mode)
            t0 = time.perf_counter()
            c2.delaunay_resident(d_p, len(pp), cfg_p)
            ms_first = (time.perf_counter() - t0) * 1e3
            sp = c2.stats()
            # later calls on the same context: device phase and host repair
            # apart, median of three, so the presizing effect shows above the
            # repair's call-to-call variation
            again, dev, rep = [], [], []
            for _ in range(3):
                t0 = time.perf_counter()
                c2.delaunay_resident(d_p, len(pp), cfg_p)
                again.append((time.perf_counter() - t0) * 1e3)
                sa = c2.stats()
                dev.append(sa["ms_device"])
                rep.append(sa["ms_repair"])
            c2.free(d_p)
            c2.close()
            presized[mode] = {"ms_first_call": round(ms_first, 3),
                              "ms_device_first_call": round(sp["ms_device"], 3),
                              "ms_repair_first_call": round(sp["ms_repair"], 3),
                              "ms_call_median": round(float(np.median(again)), 3),
                              "ms_device_median": round(float(np.median(dev)), 3),
                              "ms_repair_median": round(float(np.median(rep)), 3),
                              "value_first_call": round(len(pp) / (ms_first / 1e3), 1),
                              "slots_requested": sp["presize_slots"], "slots_allocated": sp["tet_capacity"],
                              "slots_used": sp["tet_slots"],
                              "grow_events": sp["grow_events"], "peak_active": sp["peak_active"]}
            print(f"presize {mode}: {ms_first:.1f} ms first call, device {sp['ms_device']:.1f} ms, "
                  f"later calls: device median {np.median(dev):.1f} ms", file=sys.stderr, flush=True)

    # the reference's config 1: 10,000 points U[0,1)^3 seed 0, Hilbert order,
    # exact predicates only, one instance (median of resident calls)
    cfg1 = None
    p1 = W.uniform(10_000, 0)
    c1 = wtet_amd.Config(predicate_mode="exact", enable_hilbert_sorting=True)
    if not args.no_config1:
        d_1 = ctx.upload(p1)
        for _ in range(3):
            ctx.delaunay_resident(d_1, len(p1), c1)
        t1 = []
        for _ in range(20):
            t0 = time.perf_counter()
            ctx.delaunay_resident(d_1, len(p1), c1)
            t1.append((time.perf_counter() - t0) * 1e3)
        ctx.free(d_1)
        ms1 = float(np.median(t1))
        r1 = ctx.delaunay(p1, c1)
        cfg1 = {"workload": "uniform_10000_seed0", "points": len(p1), "order": "hilbert",
                "predicates": "exact only", "ms_median": round(ms1, 3), "calls": len(t1),
                "value": round(len(p1) / (ms1 / 1e3), 1)}
        print(f"config 1 (10k, Hilbert, exact only): {ms1:.2f} ms median", file=sys.stderr, flush=True)

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle as O
        pn, _, _ = O.normalize_points(pts)
        t0 = time.perf_counter()
        tets = O.delaunay(pn)
        dt = time.perf_counter() - t0
        cpu = {"value": args.n / dt, "unit": "points/s", "cores": 1, "kind": "port",
               "sample": f"the whole {args.n}-point input, serial Bowyer-Watson of the oracle "
                         f"({dt:.1f} s, exact predicates + SoS, walk from the last new tet)",
               "host": host_info(), "threads": 1}
        ref = O.canonical_real(tets, args.n)
        got = O.canonical_real(res.tets, args.n)
        parity = bool(ref.shape == got.shape and np.array_equal(ref, got))
    if not args.no_cpu and cfg1 is not None:
        # config 1 on the CPU: the same oracle, and full_bw_reconstruction as
        # written (src/phase2.rs:383-606: input order, linear seed scan)
        pn1, _, _ = O.normalize_points(p1)
        t0 = time.perf_counter()
        ref1 = O.canonical_real(O.delaunay(pn1), len(p1))
        dt_walk = time.perf_counter() - t0
        t0 = time.perf_counter()
        ref1b = O.canonical_real(O.delaunay_reference(pn1), len(p1))
        dt_scan = time.perf_counter() - t0
        cfg1["parity_with_oracle"] = bool(np.array_equal(O.canonical_real(r1.tets, len(p1)), ref1)
                                          and np.array_equal(ref1, ref1b))
        cfg1["cpu"] = {"oracle_walk_seed_ms": round(dt_walk * 1e3, 1),
                       "reference_linear_seed_scan_ms": round(dt_scan * 1e3, 1), "cores": 1}
        cpu["config1"] = cfg1["cpu"]
    if not args.no_cpu and batch is not None:
        # the batched workload on every host core: the oracle per cloud,
        # clouds spread over OpenMP threads (each cloud normalized on its own)
        threads = min(16, os.cpu_count() or 1)
        pnb = np.concatenate([O.normalize_points(bp[int(boff[c]):int(boff[c + 1])])[0]
                              for c in range(len(boff) - 1)])
        t0 = time.perf_counter()
        O.delaunay_many(pnb, boff, threads)
        dt_b = time.perf_counter() - t0
        batch["cpu_all_cores"] = {"value": round(len(bp) / dt_b, 1), "unit": "points/s", "threads": threads,
                                  "kind": "port", "s": round(dt_b, 2),
                                  "sample": "the whole 4,096-cloud batch, one oracle Bowyer-Watson per cloud"}
        cpu["batch_all_cores"] = batch["cpu_all_cores"]
        print(f"cpu batch on {threads} threads: {dt_b:.1f} s", file=sys.stderr, flush=True)

    kern = {k: {"launches": v[0], "ms": round(v[1], 3)} for k, v in kt.items() if v[0]}
    out = {
        "metric": "points/s @1M uniform f32: plain vs filtered+batched, same GPU, bit-exact tets",
        "value": round(value, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"uniform_{args.n}_seed{args.seed}", "points": args.n,
                   "distribution": "i.i.d. U[0,1)^3 f32, numpy PCG64", "predicates": "f32->f64->exact",
                   "batched": False, "insertion_ring": args.ring, "autotuned_items_per_thread": items,
                   "flip_graphs": not args.no_graphs},
        "value_definition": "GPU loop only: points resident in HBM before the timed region, tets and "
                            "adjacency left in HBM (host repair included); end_to_end adds the "
                            "host-to-device and device-to-host copies",
        "roofline": roofline,
        "kernel_rooflines": per_kernel,
        "flip_counts": flip_counts,
        "direct_launches": direct,
        "cpu_baseline": cpu,
        "config1": cfg1,
        "plain_exact": plain,
        "speedup_vs_plain": round(value / plain["value"], 3) if plain else None,
        "end_to_end": end_to_end,
        "batched": batch,
        "stress": stress,
        "presize": presized,
        "parity_with_oracle": parity,
        "stage_fractions": stages,
        "ms_device": round(st_t["ms_device"], 3),
        "ms_repair": round(st_t["ms_repair"], 3),
        "ms_repair_staging": round(st_t["ms_repair_staging"], 3),
        "ms_repair_compute": round(st_t["ms_repair_compute"], 3),
        "repair_rows_staged": st_t["repair_rows_staged"],
        "repair_fetches": st_t["repair_fetches"],
        "violations_gpu": st["violations_gpu"],
        "repair_regions": st["repair_regions"],
        "rounds": st["rounds"],
        "flip_iterations": st["flip_iterations"],
        "job_algorithmic_bytes": job_bytes(st),
        "job_achieved_GBs": round(job_bytes(st) / (ms / 1e3) / 1e9, 1),
        "kernel_ms": kern,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
