"""Benchmark of the simulation step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain-1m]

One step is one NeuCor::run() with runAll = true and runSpeed = dt = 1 ms over
the whole network (drive, neuron update, STDP, fan-out).  The network and its
state are resident in HBM before the timed region starts.  Prints one JSON
line with the throughput in synapse-updates/s (one synapse-update = one
(synapse, step) pair), the roofline of the dominant kernel (HIP events around
each of its launches inside the timed region) and the reference's own
NeuCor::run() timed on this host's cores on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "synapse updates/s (% HBM roofline); served sims/s + p50/p99 vs plain"
HBM_PEAK_GBS = 8000.0
TIMING_EVERY = 10  # kernel events on every 10th step's launches at most; at least 5 timed launches per kernel


def timing_every(steps):
    """The spacing of the steps whose launches carry kernel events: every
    TIMING_EVERY-th step, closer on short runs so that at least 5 launches
    per kernel are timed."""
    return min(TIMING_EVERY, max(1, steps // 5))


# Launch shapes that the live tuner chose for a plain workload on MI355X, used
# by runs too short to tune (under 1000 timed steps). plain-1m: 1 slice per
# workgroup with the XCD-ordered remap, 1.367 ms per 16-step trial after the
# warm-up against 1.421 ms for the engine's default (4 slices per workgroup, no
# remap); 20 timed steps from step 6, three runs each: 73.1 against 75.6 us per
# step (an earlier build).
TUNED_LAUNCH = {"plain-1m": {"waves_per_block": 1, "xcd_swizzle": 1}}


def neuron_kernel_bytes(n_neurons, c0, c1, steps):
    """Algorithmic bytes per launch of the neuron-step kernel (k_neuron), from
    the engine's event counters over the timed steps:
      20 B per neuron        in-degree/flags + input group 8, V read+write 8,
                             lastFire read 4
      32 B per 64 neurons    active-slice record 16, pending counts 8,
                             arrival-bucket count read + reset 8
      16 B per fired-list    {mask lo, mask hi, step, slice | share} written
           record            (a share of a fired slice's out-edges)
       4 B per mask word     read (words a summary bit flags) and 4 B per
                             word written back
       8 B per active slot   (depolarisation, arrival) from the active-slot
           in the cache      cache
      16 B per active slot   slot record, for active slots past the cache
           past the cache
       8 B per bucket entry  {slot, depolarisation} of an arrival
      16 B per record read   at arrival (the target fired recently, or the
                             spike came through a pending list)
       4 B per weight        written by the arrival's pre-side STDP
      16 B per cache record  written (lanes whose active set changed)
      16 B per in-synapse    post-side STDP of a fired neuron: record read
           of a fired neuron
       4 B per weight        written by post-side STDP (bits changed)
      12 B per fired neuron  firings read+write 8, lastFire write 4
       4 B per pending entry scanned (spikes whose delays exceed NC_NB steps or
                             found their bucket full, in the list of the step's
                             window)"""
    d = {k: (c1[k] - c0[k]) / steps for k in ("active_slot_steps", "arrivals", "fired", "fired_in_degree",
                                               "pending_scanned", "mask_words_read", "mask_words_written",
                                               "bucket_entries", "records_read", "arrival_weight_writes",
                                               "cache_records_written", "arrival_records_read", "fired_records",
                                               "post_weight_writes")}
    cached = d["active_slot_steps"] - d["records_read"]
    b = (20.0 * n_neurons + 32.0 * n_neurons / 64 + 4.0 * (d["mask_words_read"] + d["mask_words_written"])
         + 8.0 * cached + 16.0 * d["records_read"] + 8.0 * d["bucket_entries"] + 16.0 * d["arrival_records_read"]
         + 4.0 * d["arrival_weight_writes"] + 16.0 * d["cache_records_written"] + 16.0 * d["fired_in_degree"]
         + 4.0 * d["post_weight_writes"] + 16.0 * d["fired_records"] + 12.0 * d["fired"] + 4.0 * d["pending_scanned"])
    return b, {"A": d["active_slot_steps"], "A_records": d["records_read"], "R": d["arrivals"],
               "R_bucket": d["bucket_entries"], "R_records": d["arrival_records_read"],
               "R_w": d["arrival_weight_writes"], "cache_w": d["cache_records_written"], "F": d["fired"],
               "F_in": d["fired_in_degree"], "F_in_w": d["post_weight_writes"], "records": d["fired_records"],
               "pending": d["pending_scanned"], "mask_words": d["mask_words_read"],
               "mask_words_written": d["mask_words_written"]}


def fanout_kernel_bytes(c0, c1, steps):
    """Algorithmic bytes per launch of the fan-out kernel (k_fanout), from the
    engine's event counters over the timed steps:
      16 B per fired-list record    {mask lo, mask hi, step, slice | share}
                                    read
       8 B per fired neuron         out-row head {first edge, out-degree}
      16 B per out-edge of a        edge record {target slot, target slice |
           fired neuron             delay, latched arrival, target slice base}
      44 B per accepted spike       target slot w, last arrival, arrival read
                                    12, depol, last arrival, arrival write 12,
                                    edge latch write 4, bucket count
                                    (read-modify-write) 8, bucket entry write 8
                                    (or pending count 8 + entry 4)"""
    d = {k: (c1[k] - c0[k]) / steps for k in ("fired_records", "fired", "fired_out_degree", "accepted_spikes")}
    b = 16.0 * d["fired_records"] + 8.0 * d["fired"] + 16.0 * d["fired_out_degree"] + 44.0 * d["accepted_spikes"]
    return b, {"records": d["fired_records"], "F": d["fired"], "F_out": d["fired_out_degree"],
               "accepted": d["accepted_spikes"]}


def copy_calibration(mib):
    """Two plain streaming kernels of known size (torch.neg over `mib` MiB of
    float32 into a second buffer, then a device-to-device copy of the same
    size: each reads and writes exactly that many bytes), five launches each,
    for scaling the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": 5}), flush=True)


def drive_fires(drive, ids, groups, n_corr_groups, steps, keep):
    """Input fires of steps 1..steps for the inputs in `keep` (the device's
    group draw restated on the host with the oracle's Philox)."""
    from oracle import oracle as O
    key = (drive.seed & 0xFFFFFFFF, drive.seed >> 32)
    per = [[] for _ in range(steps)]
    members = {}
    for i, g in zip(ids.tolist(), groups.tolist()):
        if i in keep:
            members.setdefault(g, []).append(i)
    for s in range(1, steps + 1):
        k = s % drive.period
        for g, mem in members.items():
            amp = drive.amp_corr if g < n_corr_groups else drive.amp_ind
            p = drive.p0 * (1.0 + amp * math.sin(2.0 * math.pi * k / drive.period))
            thr = 0 if p <= 0 else (0xFFFFFFFF if p >= 1 else int(math.floor(p * 4294967296.0)))
            u = int(O.philox((s, g, 4, 0), key)[0])
            if u < thr:
                per[s - 1].extend(mem)
    return [sorted(x) for x in per]


def cpu_baseline(net, drive, budget_s=12.0, sample_neurons=10_000, timed_steps=30):
    """The reference's NeuCor::run() (oracle/_ref/ref_harness, built from the
    reference's src/NeuCor.cpp) on the first `sample_neurons` neurons of the
    workload (a compact region, ids follow a Morton curve) with their
    internal synapses, `timed_steps` steps after 100 warm-up steps; single
    thread.  The whole 10^6-neuron network is not used: the reference builds
    it in O(S K) createSynapse calls and steps it in seconds per step, which
    would keep the benchmark busy for many minutes; its cost per synapse-step
    is that of any region of the same density and degree."""
    from oracle import oracle as O
    from oracle import ref as R
    ex = net.export()
    M = min(sample_neurons, len(ex["in_ptr"]) - 1)
    ptr = ex["in_ptr"][:M + 1]
    src = ex["in_src"][:ptr[M]]
    keep = src < M
    rows = np.repeat(np.arange(M), np.diff(ptr))[keep]
    sub = {"in_src": src[keep].copy(), "w": ex["w"][:ptr[M]][keep].copy(),
           "delay": ex["delay"][:ptr[M]][keep].copy(),
           "in_ptr": np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=M))]).astype(np.int64)}
    S = len(sub["in_src"])
    ids, groups = net.select_inputs(drive)
    ncorr = drive.n_groups if drive.n_correlated > 0 else 0
    cores = 1
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    if R.available():
        warm = 100

        def drive_for(total):
            # the workload's input fires and raw background draws (the
            # reference applies its own refractory check to the latter)
            inputs = drive_fires(drive, ids, groups, ncorr, total, set(range(M)))
            bg = [[] for _ in range(total)]
            if drive.bg_prob > 0:
                es, ei = O.bg_events(drive.seed, drive.bg_prob, M, 0, total)
                for s_, i_ in zip(es.tolist(), ei.tolist()):
                    bg[s_ - 1].append(i_)
            return inputs, bg

        steps = timed_steps
        inp, bg = drive_for(warm + steps)
        res = R.run_bench(sub, steps, inputs=inp, bg=bg, warmup=warm)
        return {"value": S * steps / res["run_s"], "unit": "synapse-updates/s", "cores": cores,
                "kind": "reference",
                "sample": "reference NeuCor::run() (runAll, runSpeed = 1 ms, STDP on, sign flags repaired, the "
                          "workload's input and background fires) on the first %d neurons of the workload network "
                          "(a compact Morton region) and their %d internal synapses, %d timed steps after %d "
                          "warm-up steps; 1 thread on %s (nproc %d)"
                          % (M, S, steps, warm, cpu_model, os.cpu_count() or 0)}
    o = O.Sim(sub)
    o.set_drive(ids[ids < M], groups[ids < M], ncorr, drive.period, drive.p0, drive.amp_corr, drive.amp_ind,
                0.0, drive.seed)
    t0 = time.perf_counter()
    o.step(3, cap=M * 3)
    per_step = max((time.perf_counter() - t0) / 3, 1e-6)
    steps = int(max(5, min(5000, budget_s / per_step)))
    t0 = time.perf_counter()
    o.step(steps, cap=M * steps)
    dt = time.perf_counter() - t0
    return {"value": S * steps / dt, "unit": "synapse-updates/s", "cores": cores, "kind": "port",
            "sample": "C oracle (clocked restatement of NeuCor::run) on the first %d neurons and %d internal "
                      "synapses, %d steps; 1 thread on %s" % (M, S, steps, cpu_model)}


def bench_batched(args):
    """The batched form: `args.nets` independent networks (log-uniform size,
    uneven degree, per-network seed and input correlation) packed into one
    simulation, launch shape autotuned on the live run, every network run for
    its full step budget.  Beside it, a sample of the same networks run one
    after another on the plain path (one simulation each, default launch),
    whose final state must equal the batched run's bit for bit."""
    import concurrent.futures as cf

    import neurocorrelation_amd as nc
    from neurocorrelation_amd import workloads as W

    steps = args.steps
    specs = W.batched_specs(count=args.nets, seed=0, steps=steps)
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(16) as ex:
        built = list(ex.map(W.build_spec, specs))
    sizes = [b[0].device_size() for b in built]
    n_syn = [b[0].n_syn for b in built]
    n_neu = [b[0].n for b in built]
    S_tot, N_tot = sum(n_syn), sum(n_neu)
    print("[bench] batched: %d networks, %d neurons, %d synapses, built in %.1fs" %
          (len(built), N_tot, S_tot, time.perf_counter() - t0), file=sys.stderr, flush=True)
    sim = nc.Simulation(sum(p for p, _ in sizes), S_tot, max_nets=len(built), raster_capacity=0,
                        max_slots=sum(s for _, s in sizes))
    rng = np.random.default_rng(1)
    sample = sorted(rng.choice(len(built), size=min(args.sample, len(built)), replace=False).tolist())
    ids = []
    for j, (net, drive, params, st) in enumerate(built):
        ids.append(sim.add(net, params, drive, steps=st))
        if j not in sample:
            net.close()
        if (j + 1) % 1024 == 0:
            print("[bench] added %d networks" % (j + 1), file=sys.stderr, flush=True)
    setup_s = time.perf_counter() - t0
    print("[bench] setup %.1fs" % setup_s, file=sys.stderr, flush=True)

    sim.synchronize()
    tune_table, tune_s = None, 0.0
    if args.tune_steps:
        # the engine's 38 candidate shapes, each timed from the same saved
        # state over 3 rounds (tuning takes no simulated step)
        t_tune = time.perf_counter()
        gs = (0, min(args.tune_steps, 1024))
        cands = [dict(waves_per_block=w, xcd_swizzle=x & 1, graph_steps=g, stage_tables=x >> 1)
                 for g in gs for w in (1, 2, 4, 8) for x in range(4)]
        cands += [dict(waves_per_block=w, graph_steps=g, persistent=1) for g in gs for w in (2, 4, 8)]
        chosen, ms = sim.tune_list(cands, steps_per_trial=args.tune_steps, rounds=3)
        tune_s = time.perf_counter() - t_tune
        tune_table = [dict(c, median_ms=float(m)) for c, m in zip(cands, ms)]
        for row in tune_table:
            print("[bench] tune %s" % row, file=sys.stderr, flush=True)
        print("[bench] tuned launch %s (%.2fs)" % (chosen, tune_s), file=sys.stderr, flush=True)
    else:
        sim.set_launch(args.waves_per_block)
        chosen = sim.launch()
    done = 0
    t_start = time.perf_counter()
    while done < steps:
        k = min(args.chunk, steps - done)
        sim.step(k)
        done += k
        if done % 5000 < args.chunk:
            sim.synchronize()
            print("[bench] batched step %d (%.1fs)" % (done, time.perf_counter() - t_start), file=sys.stderr,
                  flush=True)
    sim.synchronize()
    elapsed = time.perf_counter() - t_start
    assert all(sim.progress(i)[1] == 0 for i in ids)
    mine = {j: (sim.potential(ids[j], n_neu[j]), sim.weights(ids[j], n_syn[j]), sim.firings(ids[j], n_neu[j]))
            for j in sample}
    sim.close()

    # the same networks one after another on the plain path (a sample)
    seq_t = []
    for j in sample:
        net, drive, params, st = built[j]
        alone = nc.Simulation.for_network(net)
        a = alone.add(net, params, drive)
        alone.synchronize()
        t = time.perf_counter()
        left = st
        while left:
            k = min(args.chunk, left)
            alone.step(k)
            left -= k
        alone.synchronize()
        seq_t.append(time.perf_counter() - t)
        V, w, f = mine[j]
        assert np.array_equal(alone.potential(a, n_neu[j]).view(np.uint32), V.view(np.uint32)), "V differs, net %d" % j
        assert np.array_equal(alone.weights(a, n_syn[j]).view(np.uint32), w.view(np.uint32)), "w differs, net %d" % j
        assert np.array_equal(alone.firings(a, n_neu[j]), f), "firings differ, net %d" % j
        alone.close()
        net.close()
    seq_per_sim = float(np.mean(seq_t)) if seq_t else None
    sims_s = len(built) / elapsed
    out = {
        "metric": "batched: sims/s and synapse-updates/s vs sequential plain runs",
        "value": sims_s, "unit": "sims/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": elapsed * 1000.0 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched-%d" % len(built), "networks": len(built), "neurons": N_tot,
                   "synapses": S_tot, "steps_per_network": steps, "launch": chosen,
                   "tune_steps_per_trial": args.tune_steps},
        "synapse_updates_per_s": S_tot * steps / elapsed,
        "tune_s": tune_s, "sims_per_s_with_tuning": len(built) / (elapsed + tune_s), "tune_table": tune_table,
        "sequential_plain": {"sample": len(sample), "mean_s_per_sim": seq_per_sim,
                             "sims_per_s": 1.0 / seq_per_sim if seq_per_sim else None, "identical_final_state": True},
        "speedup_vs_sequential": sims_s * seq_per_sim if seq_per_sim else None,
        "setup_s": setup_s,
    }
    print(json.dumps(out), flush=True)


def bench_served(args):
    """The served form: a Poisson stream of jobs (`args.rate` per second for
    `args.window` seconds; network size, degree and step count drawn per job)
    into the native server (nc_server_*): networks built on builder threads
    (on the GPU with --device-wiring), admitted by free room and batched
    continuously, evicted and resumed under overload with --evict-after;
    completion latency from arrival (network construction included) and
    sims/s.  Beside it, a sample of the same jobs run one after another on the
    plain path (construction included), whose results must equal the served
    run's bit for bit; from the sample, a per-job time model (least squares
    over constant, synapses, synapse-steps) gives the p50/p99 latency of the
    same arrival stream served one job at a time in arrival order."""
    from scipy.optimize import nnls
    from neurocorrelation_amd import serve as SV

    rng = np.random.default_rng(7)
    n_jobs = int(round(args.rate * args.window))
    weights = (0.2, 0.8, 0.0) if args.runaway else (0.05, 0.3, 0.2)
    specs = SV.served_specs(n_jobs, seed=3, n_range=(300, args.max_neurons), k_range=(20, 100),
                            steps_range=(args.min_job_steps, args.max_job_steps), weights=weights)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=n_jobs))
    jobs = [SV.Job(i, sp, float(t)) for i, (sp, t) in enumerate(zip(specs, arrivals))]
    server = SV.Server(0, 0, max_nets=4096, chunk=args.chunk, lanes=args.lanes, builders=args.builders,
                       device_wiring=args.device_wiring, evict_after=args.evict_after,
                       overflow_lanes=args.overflow_lanes, hbm_fraction=0.5)
    print("[bench] served: %d jobs over %.0fs" % (n_jobs, args.window), file=sys.stderr, flush=True)
    done, rest = server.run(jobs, t_limit=args.wall_cap,
                            log=lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
    server.close()
    ok = [j for j in done if "digest" in j.result]
    lat = np.array([j.t_done - j.arrival for j in done])
    t_end = max(j.t_done for j in done) if done else 0.0
    syn_steps = sum(j.n_syn * j.steps for j in ok)
    # the same jobs alone, one after another, on the plain path: a sample of
    # the completed ones drawn by job index (the same jobs whatever order
    # they completed in)
    ok = sorted(ok, key=lambda j: j.idx)
    sample = sorted(rng.choice(len(ok), size=min(args.sample, len(ok)), replace=False).tolist()) if ok else []
    seq_t, feats = [], []
    for i in sample:
        job = ok[i]
        res, el = SV.run_alone(job.spec)
        assert res == job.result, "job %d differs alone and in the mix" % job.idx
        seq_t.append(el)
        print("[bench] job %d alone: %.2fs, identical" % (job.idx, el), file=sys.stderr, flush=True)
        feats.append([1.0, job.n_syn, job.n_syn * job.steps])
    seq = None
    if len(seq_t) >= 3:
        coef, _ = nnls(np.array(feats), np.array(seq_t))
        # every job of the stream (sizes from the served run; unfinished jobs
        # sized from their spec's bulk degree) served alone in arrival order
        by = {j.idx: j for j in ok}
        t_job = []
        for j in sorted(jobs, key=lambda j: j.arrival):
            syn = by[j.idx].n_syn if j.idx in by else j.spec["wiring"]["n_neurons"] * j.spec["k"]
            t_job.append(coef[0] + coef[1] * syn + coef[2] * syn * j.steps)
        arr = np.array(sorted(j.arrival for j in jobs))
        fin = np.empty_like(arr)
        c = 0.0
        for k, (a_, t_) in enumerate(zip(arr, t_job)):
            c = max(a_, c) + t_
            fin[k] = c
        slat = fin - arr
        seq = {"sample": len(sample), "mean_s_per_job": float(np.mean(seq_t)), "sims_per_s": 1.0 / float(np.mean(t_job)),
               "identical_results": True,
               "model_s": {"per_job": coef[0], "per_synapse": coef[1], "per_synapse_step": coef[2]},
               "latency_s_modelled": {"p50": float(np.percentile(slat, 50)), "p99": float(np.percentile(slat, 99))}}
    out = {
        "metric": "served: sims/s and p50/p99 completion latency vs sequential plain runs",
        "value": len(ok) / t_end if t_end else None, "unit": "sims/s", "n_gpus": 1, "steps": None,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "served", "rate_per_s": args.rate, "window_s": args.window, "jobs": n_jobs,
                   "neurons": [300, args.max_neurons], "synapses_per_neuron": [20, 100],
                   "job_steps": [args.min_job_steps, args.max_job_steps], "chunk": args.chunk, "lanes": args.lanes,
                   "builders": args.builders, "device_wiring": bool(args.device_wiring),
                   "evict_after_s": args.evict_after, "overflow_lanes": args.overflow_lanes,
                   "weights": {"w_min": weights[0], "w_max": weights[1], "inhib_frac": weights[2]},
                   "wall_cap_s": args.wall_cap},
        "completed": len(ok), "refused": len(done) - len(ok), "unfinished": rest["unfinished"],
        "latency_s": {"p50": float(np.percentile(lat, 50)) if len(lat) else None,
                      "p99": float(np.percentile(lat, 99)) if len(lat) else None,
                      "max": float(lat.max()) if len(lat) else None},
        "build_s": {"p50": float(np.percentile([j.t_built - j.arrival for j in ok], 50)) if ok else None},
        "synapse_updates_per_s": syn_steps / t_end if t_end else None,
        "sequential_plain": seq,
        "speedup_vs_sequential": ((len(ok) / t_end) / seq["sims_per_s"]) if (t_end and seq) else None,
        "serving_loop": rest["loop"],
        "mean_rate_per_step": (sum(j.result["spikes"] for j in ok) / max(1, sum(j.n * j.steps for j in ok))),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 10,000 for plain workloads, 20,000 per network for batched-*)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="plain-1m")
    ap.add_argument("--chunk", type=int, default=500)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="leave HIP events out of the timed region (for external profilers)")
    ap.add_argument("--no-counts", action="store_true",
                    help="skip the counting replay (profiler runs: every launch is then uninstrumented)")
    ap.add_argument("--nets", type=int, default=4096, help="batched-* workloads: number of networks")
    ap.add_argument("--sample", type=int, default=12, help="batched-*: networks rerun on the plain path")
    ap.add_argument("--tune-steps", type=int, default=16,
                    help="steps per autotuner trial before the timed steps (0: no autotuning; batched-*: the "
                         "--waves-per-block launch, plain: the engine's default launch or --launch)")
    ap.add_argument("--waves-per-block", type=int, default=4, help="batched-*: launch shape when not autotuning")
    ap.add_argument("--launch", default="",
                    help='plain workloads: launch shape as JSON, e.g. {"persistent": 1, "waves_per_block": 8}')
    ap.add_argument("--rate", type=float, default=200.0, help="served: job arrivals per second")
    ap.add_argument("--window", type=float, default=60.0, help="served: seconds of arrivals")
    ap.add_argument("--max-neurons", type=int, default=100_000, help="served: largest job network")
    ap.add_argument("--min-job-steps", type=int, default=1_000, help="served: shortest job")
    ap.add_argument("--max-job-steps", type=int, default=50_000, help="served: longest job")
    ap.add_argument("--wall-cap", type=float, default=900.0, help="served: stop serving after this many seconds")
    ap.add_argument("--lanes", type=int, default=2,
                    help="served: simulations the GPU alternates between (host work of one overlaps the other's chunk)")
    ap.add_argument("--builders", type=int, default=8, help="served: host threads building job networks")
    ap.add_argument("--device-wiring", action="store_true", help="served: build job networks on the GPU")
    ap.add_argument("--evict-after", type=float, default=0.0,
                    help="served: seconds a job waits for room before it evicts the job with most steps left")
    ap.add_argument("--overflow-lanes", type=int, default=0,
                    help="served: extra lanes made from free HBM while jobs wait")
    ap.add_argument("--runaway", action="store_true",
                    help="served: all-excitatory weights in [0.2, 0.8) (jobs run at p ~ 0.27 instead of ~0.0025)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a streaming copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU (N = 1): the plain path's single network is measured on one "
              "MI355X; independent networks would shard across GPUs without communication",
              file=sys.stderr)
        sys.exit(2)

    if args.workload.startswith("batched"):
        args.steps = args.steps or 20_000
        bench_batched(args)
        return
    if args.workload == "served":
        bench_served(args)
        return
    args.steps = args.steps or 10_000

    import neurocorrelation_amd as nc
    from neurocorrelation_amd import workloads as W

    t0 = time.perf_counter()
    net, drive, params, _ = W.build(args.workload, seed=0)
    info = net.info()
    N, S = info["n"], info["n_syn"]
    sim = nc.Simulation.for_network(net)
    nid = sim.add(net, params, drive)
    if args.launch:
        launch, launch_source = json.loads(args.launch), "argument"
    elif args.workload in TUNED_LAUNCH:
        launch, launch_source = dict(TUNED_LAUNCH[args.workload]), "table of tuned shapes"
    else:
        launch, launch_source = {}, "engine default"
    sim.set_launch(**launch)
    setup_s = time.perf_counter() - t0
    print("[bench] %s: %d neurons, %d synapses, setup %.1fs, launch %s" % (args.workload, N, S, setup_s,
          sim.launch()), file=sys.stderr, flush=True)

    def run(sim_, n):
        done_ = 0
        while done_ < n:
            k_ = min(args.chunk, n - done_)
            sim_.step(k_)
            done_ += k_

    run(sim, args.warmup)
    sim.synchronize()
    tuning = None
    if not args.launch and args.tune_steps and args.steps >= 1000:
        # The launch shape autotuned on the live run after the warm-up steps,
        # every candidate timed from the same saved state (no simulated step
        # is taken): slices per workgroup x the XCD-ordered remap x trace
        # tables in LDS, ordinary launches (the kernel events ride on them).
        # The tuner's 6 GB of state copies slow the next millisecond or two of
        # steps (measured: 20 timed steps right after it at 119-153 us per
        # step, against 82-84 us untuned, both with kernel events on every step):
        # only runs of 1000 steps or more are tuned.
        cands = [dict(waves_per_block=w, xcd_swizzle=x & 1, stage_tables=x >> 1) for w in (1, 2, 4, 8)
                 for x in range(4)]
        t_tune = time.perf_counter()
        chosen, ms = sim.tune_list(cands, steps_per_trial=args.tune_steps, rounds=3)
        launch = {k: chosen[k] for k in ("waves_per_block", "xcd_swizzle", "stage_tables")}
        launch_source = "tuned on the live run"
        tuning = {"s": time.perf_counter() - t_tune, "steps_per_trial": args.tune_steps, "rounds": 3,
                  "table": [dict(c, median_ms=float(m)) for c, m in zip(cands, ms)]}
        print("[bench] tuned launch %s (%.2fs; default shape %.3f ms, chosen %.3f ms per trial)" %
              (launch, tuning["s"], float(ms[cands.index(dict(waves_per_block=4, xcd_swizzle=0, stage_tables=0))]),
               float(ms.min())), file=sys.stderr, flush=True)
    every = timing_every(args.steps)
    if not args.no_kernel_timing:
        # start/stop events carried by the neuron-step and fan-out dispatches
        # of every `every`-th step: each step that carries them takes about
        # 14 us longer at plain-1m (20 timed steps with events on every step:
        # 85.6 us per step, 71.9 without), so they are spread thin
        sim.set_kernel_timing((1 << 1) | (1 << 3), every)
    t_start = time.perf_counter()
    run(sim, args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t_start
    timing = {k: sim.kernel_timing(k) for k in (1, 3)}
    launch_used = sim.launch()
    sim.set_kernel_timing(0)
    print("[bench] timed %d steps in %.3fs" % (args.steps, elapsed), file=sys.stderr, flush=True)
    sim.close()
    if not args.no_kernel_timing:
        # the same steps on a fresh simulation without the events (the step is
        # deterministic: the same activity), for what the events cost
        sim = nc.Simulation.for_network(net)
        sim.add(net, params, drive)
        sim.set_launch(**launch)
        run(sim, args.warmup)
        sim.synchronize()
        t1 = time.perf_counter()
        run(sim, args.steps)
        sim.synchronize()
        print("[bench] the same %d steps without kernel events: %.2f us/step (timed run: %.2f)" %
              (args.steps, (time.perf_counter() - t1) * 1e6 / args.steps, elapsed * 1e6 / args.steps),
              file=sys.stderr, flush=True)
        sim.close()

    # Event counts of the timed steps: the same network, drive and steps again
    # with the device counters on (the step is deterministic; counting is kept
    # out of the timed run); firing counts split inputs from the rest.
    split = {"input_fired_per_step": None, "network_fired_per_step": None}
    if args.no_counts:
        c0 = c1 = None
    else:
        sim = nc.Simulation.for_network(net)
        cid = sim.add(net, params, drive)
        sim.set_launch(**launch)
        run(sim, args.warmup)
        sim.set_counting(True)
        c0 = sim.counters()
        f0 = sim.firings(cid, N).astype(np.int64)
        run(sim, args.steps)
        sim.synchronize()
        c1 = sim.counters()
        df = sim.firings(cid, N).astype(np.int64) - f0
        sim.close()
        sel, _ = net.select_inputs(drive)
        is_in = np.zeros(N, bool)
        is_in[sel] = True
        split = {"input_fired_per_step": float(df[is_in].sum()) / args.steps,
                 "network_fired_per_step": float(df[~is_in].sum()) / args.steps}

    value = S * args.steps / elapsed
    kernels = {}

    def step_bytes():
        # the persistent form's one launch per step: both phases
        b1, u1 = neuron_kernel_bytes(N, c0, c1, args.steps)
        b2, u2 = fanout_kernel_bytes(c0, c1, args.steps)
        return b1 + b2, dict(u1, **{"records": u2["records"], "F_out": u2["F_out"], "accepted": u2["accepted"]})

    models = (((1, "k_step", step_bytes),) if launch_used.get("persistent") else
              ((1, "k_neuron", lambda: neuron_kernel_bytes(N, c0, c1, args.steps)),
               (3, "k_fanout", lambda: fanout_kernel_bytes(c0, c1, args.steps))))
    for kind, name, model in models:
        launches, k_ms = timing[kind]
        b, units = model() if c0 is not None else (None, None)
        avg = (k_ms / 1000.0) / launches if launches else None
        ach = b / avg / 1e9 if (avg and b) else None
        kernels[name] = {"algorithmic_bytes_per_launch": b, "units_per_step": units,
                         "avg_launch_us": avg * 1e6 if avg else None, "launches": launches,
                         "achieved": ach, "frac": ach / HBM_PEAK_GBS if ach else None}
    first = "k_step" if "k_step" in kernels else "k_neuron"
    ev = kernels[first]["units_per_step"] or {k: None for k in ("A", "R", "F", "F_in", "pending")}
    # the dominant kernel: the larger share of the step's device time
    dom = max(kernels, key=lambda k: kernels[k]["avg_launch_us"] or 0.0)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("%s_bytes_per_launch" % dom)
    rate = ev["F"] / N if ev["F"] is not None else None

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(net, drive)
        except Exception as e:  # reported, never fatal for the GPU number
            print("[bench] cpu baseline failed: %r" % (e,), file=sys.stderr)
    K = kernels[dom]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "synapse-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict({"workload": args.workload, "neurons": N, "synapses": S, "dt_ms": params.dt,
                        "stdp": True, "seed": 0, "fired_per_step": ev["F"], "rate_per_step": rate,
                        "arrivals_per_step": ev["R"], "active_slots_per_step": ev["A"],
                        "pending_per_step": ev["pending"], "drive_p0": drive.p0, "bg_prob": drive.bg_prob,
                        "launch": launch_used},
                       **split),
        "roofline": {"bound": "hbm", "achieved": K["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": K["frac"], "traffic": traffic, "kernel": dom,
                     "algorithmic_bytes_per_launch": K["algorithmic_bytes_per_launch"],
                     "avg_launch_us": K["avg_launch_us"], "launches": K["launches"]},
        "kernels": kernels,
        "cpu_baseline": cpu,
        "launch_source": launch_source,
        "launch_tuning": tuning,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
