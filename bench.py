#!/usr/bin/env python3
"""ORB build benchmark on one MI355X.

Workload: 2^27 float32 particles from the reference generator (xorshf96,
seeds of init.cu:11) bisected to 2^10 leaves, full tree (10 levels).  A step
is one whole build from the same initial particle order, inputs resident in
HBM.  Both build forms run: the plain host-driven form (one count launch per
bisection pass per level, counts read back each pass) and the batched form
(device-driven, launch shapes autotuned before the timed builds).  `value`
is the faster form's build time.

Prints one JSON line.  `roofline` is the count-pass kernel measured with HIP
events on its own stream over the timed builds; `cpu_baseline` is the CPU
oracle (a port of the reference's o=0 path) on a bounded sample, on all the
process's host threads and on one.

Other workloads (--config): c3 clustered 2^26 -> 2^14, c4 uniform 2^28 ->
2^20 with per-level times, c5 serving (continuous batching of requests).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-load-balance_amd"))

HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TM_KEYS = ("count_ms", "partition_ms", "other_ms", "count_bytes", "count_launches", "swap_ms",
           "swap_launches", "swap_pairs", "swap_bytes")


def timed_builds(ctx, form, levels, steps):
    times, info = [], None
    for _ in range(steps):
        ctx.restore_particles()          # device copy + range reset, synchronized
        t0 = time.perf_counter()
        info = ctx.build(form, levels)   # synchronizes the stream before returning
        t1 = time.perf_counter()
        times.append((t1 - t0) * 1e3)
    return times, info


def run_form(orb, ctx, form, levels, steps, warmup, instrumented_steps):
    """Times `steps` builds with no instrumentation (the batched form replays its
    captured graph).  Then `instrumented_steps` more timed builds run with HIP
    events recorded on the library's stream around the kernels (the plain form's
    count launches: the count-pass roofline is measured live in these; the
    swap launches of either form); their own mean time is reported beside."""
    for _ in range(warmup):
        ctx.restore_particles()
        ctx.build(form, levels)
    ctx.enable_timing(False)
    times, info = timed_builds(ctx, form, levels, steps)
    ctx.enable_timing(True)
    tm_total = {k: 0 for k in TM_KEYS}
    itimes = []
    for _ in range(instrumented_steps):
        t, _ = timed_builds(ctx, form, levels, 1)
        itimes += t
        tm = ctx.timing()
        for k in TM_KEYS:
            tm_total[k] += tm[k]
    ctx.enable_timing(False)
    tm_total["instrumented_ms"] = sum(itimes) / max(len(itimes), 1)
    tm_total["instrumented_steps"] = instrumented_steps
    return times, tm_total, info


def committed_traffic(kernel="orb::k_count"):
    """HBM bytes per launch of a kernel from the committed FETCH_SIZE /
    WRITE_SIZE summary (rocprofv3 counter passes, scaled by the copy-kernel
    calibration stated in that file); None if absent."""
    path = os.path.join(ROOT, "profiles", "orb_build_c2_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == f"`{kernel}`":
            read_mb = float(cells[3].replace(",", ""))
            write_mb = float(cells[5].replace(",", ""))
            return int((read_mb + write_mb) * 1e6)
    return None


def host_threads():
    """Threads for the all-cores CPU baseline: the process's CPU share,
    bounded by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(n, d, levels, thread_counts):
    """The oracle (a port of the reference's o=0 path) building the same
    workload on the host: wall ms per thread count, with its count and
    partition phases (the reference's CountCopy and Partition columns)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    x, y, z = pyoracle.generate(n)
    out = {}
    for nt in thread_counts:
        t0 = time.perf_counter()
        r = pyoracle.build(x, y, z, d, levels_mode=levels, threads=nt)
        t1 = time.perf_counter()
        assert r.rc == 0
        out[nt] = {"ms": (t1 - t0) * 1e3, "count_ms": r.count_ms, "partition_ms": r.partition_ms}
    return out


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    k = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[k]


def serve_bench(args):
    """Serving workload: requests with N log-uniform in [2^16, 2^24] and d
    log-uniform in [2^4, 2^12], half uniform (the reference generator with the
    request's seed) and half Plummer-clustered (64 spheres, a = 0.01, 10%
    background, the request's seed), particles resident in HBM.  Run 1 (saturated): every request queued at
    t=0, builds/s = requests / makespan.  Run 2 (Poisson): arrivals at 70% of
    the saturated rate, latency = done - arrival (p50/p99).  The comparator is
    the same requests built one after another with the plain form (measured
    build + tree copy-back per request), fed the same arrival times."""
    import numpy as np
    import inputs
    import orb_amd as orb

    K = args.requests
    rng = np.random.default_rng(args.seed)
    reqs = []
    for k in range(K):
        n = int(round(2.0 ** rng.uniform(16, 24)))
        d = int(round(2.0 ** rng.uniform(4, 12)))
        kind = k % 2
        reqs.append((kind, int(rng.integers(1, 1 << 40)), n, d))
    tot_particles = sum(r[2] for r in reqs)
    log(f"[serve] {K} requests, {tot_particles} particles in all: generating their inputs")
    host = []
    for kind, s, n, d in reqs:
        if kind == 0:
            host.append(np.stack(orb.generate(n, seed=s)))
        else:
            host.append(np.stack(inputs.clustered(n, n_spheres=64, a=0.01, background=0.1,
                                                  seed=s)))

    # comparator: serial plain builds, inputs in HBM, tree copied back
    serial_plain, serial_batched = [], []
    for k, (kind, s, n, d) in enumerate(reqs):
        src = host[k]
        ctx = orb.Context(n, d, generate=False)
        ctx.set_particles(src[0], src[1], src[2])
        ctx.save_particles()
        for form, dst in ((orb.FORM_PLAIN, serial_plain), (orb.FORM_BATCHED, serial_batched)):
            ctx.restore_particles()
            t0 = time.perf_counter()
            ctx.build(form, orb.LEVELS_FULL)
            ctx.heap()
            ctx.ranges()
            dst.append((time.perf_counter() - t0) * 1e3)
        ctx.close()
    log(f"[serve] serial plain {sum(serial_plain):.1f} ms, serial batched "
        f"{sum(serial_batched):.1f} ms for all requests")

    dev = [None] * K

    def upload():
        for k in range(K):
            if dev[k] is None:
                dev[k] = [orb.DeviceArray(host[k][a]) for a in range(3)]
            else:
                for a in range(3):
                    dev[k][a].upload(host[k][a])
        orb.hip().hipDeviceSynchronize()

    def run_server(arrivals):
        upload()
        srv = orb.Server(particle_capacity=args.serve_particles, cell_capacity=1 << 20,
                         max_active=args.serve_active, levels_mode=orb.LEVELS_FULL)
        # one untimed request warms the kernels' first launches
        w = dev[0]
        wid = srv.submit_device(w[0].address, w[1].address, w[2].address,
                                reqs[0][2], reqs[0][3])
        while srv.step():
            pass
        srv.fetch(wid, reqs[0][2], reqs[0][3], particles=False)
        upload()
        ids = {}
        lat = [None] * K
        errors = 0
        t0 = time.perf_counter()
        i = done = 0
        while done < K:
            now = time.perf_counter() - t0
            while i < K and arrivals[i] <= now:
                t = dev[i]
                ids[srv.submit_device(t[0].address, t[1].address, t[2].address,
                                      reqs[i][2], reqs[i][3])] = i
                i += 1
            left = srv.step()
            while True:
                rid = srv.next_done()
                if rid is None:
                    break
                k = ids.pop(rid)
                res = srv.fetch(rid, reqs[k][2], reqs[k][3], particles=False)
                errors += res["error"] != 0
                lat[k] = res["done_ms"] * 1e-3 - (t0 + arrivals[k])
                done += 1
            if left == 0 and i < K:
                time.sleep(max(0.0, arrivals[i] - (time.perf_counter() - t0)))
        stats = srv.stats()
        srv.close()
        makespan = max(arrivals[k] + lat[k] for k in range(K)) - arrivals[0]
        return [v * 1e3 for v in lat], makespan, stats, errors

    def simulate(service_ms, arrivals):
        t, lat = 0.0, []
        for a, s in zip(arrivals, service_ms):
            t = max(t, a) + s * 1e-3
            lat.append((t - a) * 1e3)
        return lat, t - arrivals[0]

    log("[serve] saturated run")
    zero = [0.0] * K
    lat_s, span_s, st_s, err_s = run_server(zero)
    mu = K / span_s
    rate = args.rate if args.rate > 0 else 0.7 * mu
    gaps = rng.exponential(1.0 / rate, K)
    arrivals = list(np.cumsum(gaps) - gaps[0])
    log(f"[serve] Poisson run at {rate:.1f} requests/s")
    lat_p, span_p, st_p, err_p = run_server(arrivals)
    sp_lat, sp_span = simulate(serial_plain, arrivals)
    sb_lat, sb_span = simulate(serial_batched, arrivals)
    out = {
        "config": "c5", "workload": "serving: N log-uniform [2^16,2^24], d log-uniform "
                                    "[2^4,2^12], 50% uniform (xorshf96, request seed) / 50% "
                                    "Plummer (64 spheres, a=0.01, 10% background, request seed), "
                                    "Poisson arrivals, full trees, particles resident in HBM",
        "requests": K, "particles": tot_particles,
        "saturated": {"server_builds_per_s": round(mu, 2),
                      "serial_plain_builds_per_s": round(K / (sum(serial_plain) * 1e-3), 2),
                      "serial_batched_builds_per_s": round(K / (sum(serial_batched) * 1e-3), 2),
                      "server_vs_serial_plain": round(mu * sum(serial_plain) * 1e-3 / K, 3),
                      "max_batch": st_s["max_batch"], "steps": st_s["steps"], "errors": err_s},
        "poisson": {"rate_per_s": round(rate, 2),
                    "server_p50_ms": round(pct(lat_p, 50), 3),
                    "server_p99_ms": round(pct(lat_p, 99), 3),
                    "server_builds_per_s": round(K / span_p, 2),
                    "serial_plain_p50_ms": round(pct(sp_lat, 50), 3),
                    "serial_plain_p99_ms": round(pct(sp_lat, 99), 3),
                    "serial_plain_builds_per_s": round(K / sp_span, 2),
                    "serial_batched_p50_ms": round(pct(sb_lat, 50), 3),
                    "serial_batched_p99_ms": round(pct(sb_lat, 99), 3),
                    "max_batch": st_p["max_batch"], "errors": err_p},
        "note": "serial_* latencies: the measured per-request build times (inputs in HBM, "
                "tree copied back) queued first-come first-served on the same arrival times",
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--d", type=int, default=1024)
    ap.add_argument("--levels", choices=["full", "reference"], default="full")
    ap.add_argument("--forms", default="plain,batched")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", choices=["c2", "c3", "c4", "c5"], default="c2",
                    help="c2: uniform 2^27 -> 2^10; c3: clustered 2^26 -> 2^14; "
                         "c4: uniform 2^28 -> 2^20 with per-level times; c5: serving")
    ap.add_argument("--per-level", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched form with the default launch shapes")
    ap.add_argument("--tune-table", default=None,
                    help="persisted tuning table: applied when the file exists (no autotuning), "
                         "else written after autotuning")
    ap.add_argument("--requests", type=int, default=160)
    ap.add_argument("--rate", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--serve-particles", type=int, default=1 << 28)
    ap.add_argument("--serve-active", type=int, default=64)
    args = ap.parse_args()
    if args.config == "c3":
        args.log2n, args.d = 26, 1 << 14
    elif args.config == "c4":
        args.log2n, args.d, args.per_level = 28, 1 << 20, True

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)

    if args.config == "c5":
        serve_bench(args)
        return

    import orb_amd as orb
    levels = orb.LEVELS_FULL if args.levels == "full" else orb.LEVELS_REFERENCE
    n = 1 << args.log2n
    log(f"[bench] generating {n} particles ({args.config})")
    if args.config == "c3":
        import inputs
        x, y, z = inputs.clustered(n, n_spheres=64, a=0.01, background=0.1, seed=2022)
        ctx = orb.Context(n, args.d, generate=False)
        ctx.set_particles(x, y, z)
        del x, y, z
    else:
        ctx = orb.Context(n, args.d, generate=True)
    ctx.save_particles()

    res = {}
    for name in args.forms.split(","):
        form = orb.FORM_PLAIN if name == "plain" else orb.FORM_BATCHED
        log(f"[bench] form={name} warmup={args.warmup} steps={args.steps}")
        tuning = None
        if name == "batched" and args.tune_table and os.path.exists(args.tune_table):
            matched = ctx.set_tune_rows(orb.read_tune_table(args.tune_table))
            tuning = ctx.tuning()
            tuning["table"] = {"path": args.tune_table, "levels_matched": matched}
            log(f"[bench] tuning table {args.tune_table}: {matched} levels matched")
        elif name == "batched" and not args.no_autotune:
            ctx.restore_particles()
            tuning = ctx.autotune(levels, repeats=2)
            log(f"[bench] autotune: default {tuning['default_ms']:.3f} ms -> tuned "
                f"{tuning['tuned_ms']:.3f} ms (device time per build), subtree_max_keys="
                f"{tuning['subtree_max_keys']}")
            if args.tune_table:
                orb.write_tune_table(args.tune_table, ctx.tune_rows())
                log(f"[bench] tuning table written to {args.tune_table}")
        try:
            times, tm, info = run_form(orb, ctx, form, levels, args.steps, args.warmup,
                                       instrumented_steps=(args.steps if name == "plain" else 1))
        except orb.OrbError as e:
            if e.code == orb.ERR_STATE:
                log(f"[bench] form={name} unavailable")
                continue
            raise
        res[name] = {"ms": sum(times) / len(times), "min_ms": min(times), "tm": tm,
                     "passes": info["total_passes"], "level_passes": info["level_passes"]}
        if tuning is not None:
            res[name]["tuning"] = tuning
        if args.per_level:
            ctx.enable_timing(True)
            ctx.restore_particles()
            ctx.build(form, levels)
            res[name]["level_ms"] = [round(v, 4) for v in ctx.level_timing()]
            ctx.enable_timing(False)
        log(f"[bench] form={name} mean {res[name]['ms']:.3f} ms min {min(times):.3f} ms "
            f"passes={info['total_passes']} (instrumented build {tm['instrumented_ms']:.3f} ms: "
            f"count {tm['count_ms'] / tm['instrumented_steps']:.3f} ms, partition "
            f"{tm['partition_ms'] / tm['instrumented_steps']:.3f} ms, swap "
            f"{tm['swap_ms'] / tm['instrumented_steps']:.3f} ms)")
    # this GPU's own streaming bandwidth over the same arrays, to read the
    # roofline fractions against what HBM delivers in practice
    stream_read, stream_copy = ctx.measure_stream(10)
    log(f"[bench] stream over the particle arrays: read {stream_read:.0f} GB/s, "
        f"copy {stream_copy:.0f} GB/s (read + write)")
    ctx.close()

    if args.config != "c2":
        # parity-test configurations: report each form's build time and levels
        out = {"config": args.config, "n_particles": n, "leaves": args.d, "levels": args.levels,
               "forms": {k: {kk: vv for kk, vv in v.items() if kk != "tm"} for k, v in res.items()},
               "note": "level_ms: device time per level (HIP events, one instrumented build); "
                       "min-traffic GB/s per level = 28 B x n / level time"}
        if args.per_level:
            for k, v in res.items():
                if "level_ms" in v:
                    v["level_GBps_min_traffic"] = [round(28.0 * n / (ms * 1e-3) / 1e9, 1)
                                                   if ms > 0 else None for ms in v["level_ms"]]
            out["forms"] = {k: {kk: vv for kk, vv in v.items() if kk != "tm"}
                            for k, v in res.items()}
        if "plain" in res and "batched" in res:
            out["batched_speedup"] = round(res["plain"]["ms"] / res["batched"]["ms"], 3)
        print(json.dumps(out), flush=True)
        return

    best = "batched" if "batched" in res else "plain"
    r = res[best]
    plain = res.get("plain")
    tm = (plain or r)["tm"]
    achieved = tm["count_bytes"] / (tm["count_ms"] * 1e-3) / 1e9 if tm["count_ms"] > 0 else 0.0
    out = {
        "metric": "ORB build ms (2^27 pts→2^10 leaves), count-pass HBM GB/s, batched vs plain",
        "value": round(r["ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: reference xorshf96 generator (init.cu:11-25), uniform in [-0.5,0.5]^3",
        "config": {"workload": f"orb_build uniform 2^{args.log2n} -> {args.d} leaves "
                               f"({args.levels} tree)",
                   "n_particles": n, "leaves": args.d, "levels": args.levels, "form": best},
        "plain_ms": round(plain["ms"], 4) if plain else None,
        "batched_ms": round(res["batched"]["ms"], 4) if "batched" in res else None,
        "batched_speedup": (round(plain["ms"] / res["batched"]["ms"], 3)
                            if plain and "batched" in res else None),
        "passes": r["passes"],
        "autotune": res.get("batched", {}).get("tuning"),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_count (plain count-left pass)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": (committed_traffic() if (args.log2n, args.d, args.levels) == (27, 1024, "full")
                        else None),
            "traffic_source": "profiles/orb_build_c2_pmc_traffic.md (FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
            "launches": tm["count_launches"],
            "avg_launch_us": round(tm["count_ms"] * 1e3 / max(tm["count_launches"], 1), 2),
            "bytes_per_launch": int(tm["count_bytes"] / max(tm["count_launches"], 1)),
            "timed_in": f"{tm['instrumented_steps']} timed plain builds with HIP events around "
                        f"each count launch (mean {tm['instrumented_ms']:.3f} ms; plain_ms is "
                        f"timed without events)",
            "measured_read_GBps": round(stream_read, 1),
            "frac_of_measured_read": round(achieved / stream_read, 4) if stream_read > 0 else None,
        },
    }
    # the batched build's dominant kernel: the Hoare swap, one launch per level
    tb = res.get("batched", {}).get("tm")
    if tb and tb["swap_launches"] > 0:
        # algorithmic bytes: every swapped pair's two rows read and written,
        # 12 B each (x, y, z) -- 48 B x K, with K (the pairs, summed over the
        # cells of every level) counted on the device in the timed build.
        # Line bytes: at the 50% stopper density of uniform keys every line
        # of x, y, z is touched, 24 B per particle per launch read + written.
        sw_launches = tb["swap_launches"]
        sw_alg = tb["swap_bytes"] / sw_launches
        sw_line = 24 * n
        sw_us = tb["swap_ms"] * 1e3 / sw_launches
        sw_ach = sw_alg / (sw_us * 1e-6) / 1e9
        traffic = (committed_traffic("orb::k_hoare_swap<true, 3>")
                   if (args.log2n, args.d, args.levels) == (27, 1024, "full") else None)
        out["roofline_partition"] = {
            "bound": "hbm",
            "kernel": "k_hoare_swap (batched build)",
            "achieved": round(sw_ach, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(sw_ach / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_over_algorithmic": round(traffic / sw_alg, 3) if traffic else None,
            "launches": sw_launches,
            "avg_launch_us": round(sw_us, 2),
            "bytes_per_launch": int(sw_alg),
            "pairs_per_launch": int(tb["swap_pairs"] / sw_launches),
            "line_bytes_per_launch": sw_line,
            "line_GBps": round(sw_line / (sw_us * 1e-6) / 1e9, 1),
            "timed_in": "one batched build with HIP events around each swap launch, after the "
                        "timed (graph-replayed, uninstrumented) builds",
            "measured_copy_GBps": round(stream_copy, 1),
        }
    if not args.no_cpu:
        nt = host_threads()
        log(f"[bench] cpu baseline: oracle build of the same 2^{args.log2n} particles, {nt} "
            f"threads, then 1 thread")
        cb = cpu_baseline(n, args.d, levels, (nt, 1))
        mt, st = cb[nt], cb[1]
        out["cpu_baseline"] = {
            "value": round(mt["ms"], 1), "unit": "ms", "cores": nt, "kind": "port",
            "sample": f"the oracle (o=0 port, OpenMP over cells and over a large cell's keys; "
                      f"same result as 1 thread) building this workload in full (2^{args.log2n} "
                      f"particles -> {args.d} leaves) on {nt} host threads ({cpu_model()}, "
                      f"{os.cpu_count()} logical CPUs visible)",
            "count_ms": round(mt["count_ms"], 1), "partition_ms": round(mt["partition_ms"], 1),
        }
        out["cpu_baseline_1_thread"] = {
            "value": round(st["ms"], 1), "unit": "ms", "cores": 1, "kind": "port",
            "sample": "the same build on 1 thread",
            "count_ms": round(st["count_ms"], 1), "partition_ms": round(st["partition_ms"], 1),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
