#!/usr/bin/env python3
"""Headline benchmark: synaptic events/s and ms per simulated second of one
N=100,000, Nsyn=1000 Izhikevich network (balanced regime) on one MI355X.

A "step" is one 1 ms simulation step of the whole network: Izhikevich update
of every neuron, spike compaction, propagation of the step's spikes.  The
network (1e8 synapses, 800 MB) is resident in HBM before the timed region.

Prints ONE JSON line (rank 0).  Extra lines (per-scheme sweep) go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
CPU_CAP = 16  # threads of the CPU baseline: one GPU's share of the host

# dominant kernel of each (scheme, mode): the one launch that carries a step
KERNEL_OF = {("SKL", "atomic"): "k_skl_slice", ("AB", "atomic"): "k_slice_step", ("AP", "atomic"): "k_ap",
             ("N", "atomic"): "k_push_rows", ("S", "atomic"): "k_push_flat",
             ("SKL", "det"): "k_skl_slice", ("AB", "det"): "k_slice_step", ("N", "det"): "k_slice_step",
             ("S", "det"): "k_slice_step", ("AP", "det"): "k_slice_step"}


def step_bytes(n, nsyn, spikes):
    """Algorithmic bytes: 40 B per neuron-step (read v,u,a,b,c,d,Isyn; write
    v,u,Isyn) + per spike 4 B queue entry + 8 B per synaptic event."""
    return 40.0 * n + (4.0 + 8.0 * nsyn) * spikes


def profiled_traffic(kernel, alg, accum, n, nsyn, regime, slices, block):
    """The committed rocprofv3 FETCH_SIZE/WRITE_SIZE record of `kernel` on the
    same workload and plan (profiles/traffic.json), or None."""
    return traffic_record(f"{kernel}|{alg}|{accum}|n={n}|nsyn={nsyn}|regime={regime}|slices={slices}|block={block}")


def traffic_record(key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def traffic_for_window(rec, alg_bytes_per_step, steps_per_launch):
    """HBM bytes per launch of this window from a profiled record: the
    profiled run's measured bytes per step over its algorithmic bytes per
    step (which carry its own spikes per step), times this window's
    algorithmic bytes per step and steps per launch."""
    if not rec or "bytes_per_step" not in rec or "alg_bytes_per_step" not in rec:
        return None, None
    ratio = rec["bytes_per_step"] / rec["alg_bytes_per_step"]
    return ratio * alg_bytes_per_step * steps_per_launch, ratio


def kernel_bytes(alg, accum, n, nsyn, spikes):
    """Algorithmic bytes of the dominant kernel for a step with `spikes`
    spikes: the whole step, except for N/S atomic whose separate update
    kernel carries the 40 B per neuron."""
    if accum == "atomic" and alg in ("N", "S"):
        return (4.0 + 8.0 * nsyn) * spikes
    return step_bytes(n, nsyn, spikes)


def cpu_baseline(n, nsyn, seed, regime, steps):
    """The CPU oracle (fp32, -O2 -ffp-contract=off) on the first `steps`
    steps of the same network, connectivity build excluded: on all host
    cores (at most 16, the host share of one GPU) -- each thread owns a
    range of neurons and of targets, results bit-identical to one thread --
    and on one core."""
    from oracle import pyoracle
    host = os.cpu_count() or 1
    cores = max(1, min(CPU_CAP, host))
    net = pyoracle.Network(n, nsyn, seed, regime=regime)
    t0 = time.perf_counter()
    _, counts, _ = net.run(steps, raster=False, threads=cores)
    dt = time.perf_counter() - t0
    one_steps = max(1, steps // 4)
    net1 = pyoracle.Network(n, nsyn, seed, regime=regime)
    t1 = time.perf_counter()
    _, c1, _ = net1.run(one_steps, raster=False)
    dt1 = time.perf_counter() - t1
    ev = float(counts.sum()) * nsyn
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": ev / dt, "unit": "synaptic events/s", "cores": cores, "kind": "port",
            "host_cores": host, "cpu_model": model,
            "cap": f"at most {CPU_CAP} threads: the host share of one GPU on the benchmark machine",
            "sample": f"first {steps} steps of the same N={n} Nsyn={nsyn} network ({int(counts.sum())} spikes, "
                      f"{dt:.2f} s on {cores} threads of the C oracle)",
            "ms_per_sim_second": dt * 1000.0 / steps * 1000.0,
            "single_core": {"value": float(c1.sum()) * nsyn / dt1, "ms_per_sim_second": dt1 * 1e6 / one_steps,
                            "sample": f"first {one_steps} steps, 1 thread ({dt1:.2f} s)"}}


def batch_specs(count, seed):
    """Instances of the batched workload: N log-uniform on [1e3, 5e4], Nsyn from {100, 500,
    1000}, independent seeds, thalamic scale U(0.8, 1.5)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(float(np.exp(rng.uniform(np.log(1000.0), np.log(50000.0))))))
        nsyn = min(int(rng.choice([100, 500, 1000])), n - 1)
        out.append((n, nsyn, seed * 100003 + i, float(rng.uniform(0.8, 1.5))))
    return out


def bench_batch(g, args):
    specs = batch_specs(args.batch, args.seed)
    total = args.warmup + args.steps
    tuned = None
    if args.tune and not args.slice_len and args.schedule == "step":
        # the batch autotuner on this very mix (same seeds), before the timed
        # batch is built; its per-bucket slice lengths are then used by add()
        t_tune = time.perf_counter()
        tuned = g.batch_tune_mix(specs, accum=args.accum, trial_steps=400, max_trials=args.tune_trials,
                                 chunk=args.chunk, streams=args.streams, block=args.batch_block)
        tuned["tune_s"] = time.perf_counter() - t_tune
        print(f"mix tuner: {tuned}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = g.Batch(args.accum, chunk=args.chunk, slice_len=args.slice_len, schedule=args.schedule,
                block=args.batch_block, streams=args.streams)
    slots = [b.add(n, ns, sd, total, thal_scale=sc, keep_raster=False) for (n, ns, sd, sc) in specs]
    print(f"batch of {len(slots)} built in {time.perf_counter() - t0:.1f} s, "
          f"{sum(n for n, _, _, _ in specs)} neurons, {sum(n * ns for n, ns, _, _ in specs) / 1e9:.2f}e9 synapses",
          file=sys.stderr, flush=True)
    done = 0
    while done < args.warmup:
        b.step()
        done += args.chunk
    c0 = [len(b.counts(s)) for s in slots]
    t1 = time.perf_counter()
    dev_ms, launches = 0.0, 0
    while b.step():
        tm = b.timing()
        dev_ms += tm["device_ms"]
        launches += tm["launches"]
    wall = time.perf_counter() - t1
    spikes = events = 0
    byts = 0.0
    for s, (n, ns, _, _) in zip(slots, specs):
        c = b.counts(s)[c0[slots.index(s)]:]
        spikes += int(c.sum())
        events += int(c.sum()) * ns
        byts += 40.0 * n * len(c) + (4.0 + 8.0 * ns) * float(c.sum())
    steps_timed = total - c0[0]
    b.close()
    sim_s = len(specs) * steps_timed / 1000.0
    # back-to-back plain runs of a sample, one network per handle (AB scheme)
    idx = list(range(0, len(specs), max(1, len(specs) // max(1, args.plain_sample))))[:args.plain_sample]
    p_wall, p_sim = 0.0, 0.0
    for i in idx:
        n, ns, sd, sc = specs[i]
        with g.Network(n, ns, sd, thal_scale=sc) as net:
            net.run(args.warmup, alg="AB", accum=args.accum, raster=False)
            t2 = time.perf_counter()
            net.run(args.steps, alg="AB", accum=args.accum, raster=False)
            p_wall += time.perf_counter() - t2
            p_sim += args.steps / 1000.0
    achieved = byts / (dev_ms / 1e3) / 1e9 if dev_ms > 0 else 0.0
    return {
        "metric": "aggregate simulated seconds per second, batched networks",
        "value": sim_s / wall, "unit": "sim-s/s", "n_gpus": 1, "steps": steps_timed, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps_timed, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Philox networks)",
        "config": {"workload": f"{len(specs)} networks, N log-uniform [1e3, 5e4], Nsyn in {{100,500,1000}}, "
                               f"thalamic x U(0.8,1.5), {steps_timed} timed steps each",
                   "accum": args.accum, "chunk": args.chunk,
                   "slice_len": args.slice_len or ("tuned per bucket on the mix" if tuned else "min(768, N)"),
                   "tuned": tuned,
                   "schedule": args.schedule, "block": args.batch_block or 256, "streams": args.streams or 3,
                   "neurons": sum(x[0] for x in specs),
                   "synapses": sum(x[0] * x[1] for x in specs), "seed": args.seed},
        "events_per_s": events / wall,
        "plain": {"sim_s_per_s": p_sim / p_wall, "sample": len(idx), "scheme": "AB"},
        "speedup_vs_plain": (sim_s / wall) / (p_sim / p_wall),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     # profiled at the untuned plan (768-neuron slices); the mix
                     # tuner moves one or two of 55 buckets of this mix
                     "traffic": (traffic_record(f"k_batch_step|batch{len(specs)}|{args.accum}|chunk={args.chunk}"
                                                f"|slice_len=768|streams={args.streams or 3}|per_step") or {})
                     .get("bytes_per_step")
                     if not args.slice_len and args.schedule == "step" and args.seed == 20221 else None,
                     "traffic_plan": "768-neuron slices, untuned",
                     "kernel": "k_batch_step",
                     # each step is one launch per instance group, the groups'
                     # launches running concurrently on their own streams:
                     # bytes and device time are per step of the whole batch
                     "bytes_per_step": byts / max(1, steps_timed),
                     "device_us_per_step": dev_ms * 1e3 / max(1, steps_timed),
                     "launches_per_step": launches / max(1, steps_timed)},
    }


def bench_serve(g, args):
    """Open-loop serving: Poisson arrivals for args.serve seconds."""
    import numpy as np
    rng = np.random.default_rng(args.seed + 1)
    t_arr, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / args.rate))
        if t >= args.serve:
            break
        t_arr.append(t)
    specs = batch_specs(len(t_arr), args.seed + 2)
    steps = [int(x) for x in rng.integers(100, 5001, len(t_arr))]
    srv = g.Server(args.accum, chunk=args.chunk, budget=int(args.budget_gb * (1 << 30)), slice_len=args.slice_len)
    ids = []
    t0 = time.perf_counter()
    last_report = t0
    for k, ta in enumerate(t_arr):
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        n, ns, sd, sc = specs[k]
        ids.append(srv.submit(n, ns, sd, steps[k], thal_scale=sc, keep_raster=False))
        if time.perf_counter() - last_report > 10:
            st = srv.stats()
            print(f"t={time.perf_counter() - t0:.0f}s submitted {st['submitted']} done {st['completed']} "
                  f"active {st['active']} pending {st['pending']}", file=sys.stderr, flush=True)
            last_report = time.perf_counter()
    srv.drain(600)
    wall = time.perf_counter() - t0
    st = srv.stats()
    infos = [srv.info(r) for r in ids]
    lat = np.array([i["done_s"] - i["submit_s"] for i in infos])
    first = min(i["submit_s"] for i in infos)
    last = max(i["done_s"] for i in infos)
    sim_s = sum(steps) / 1000.0
    events = sum(i["spikes"] * specs[k][1] for k, i in enumerate(infos))
    srv.close()
    # back-to-back plain runs of a sample of the same requests
    idx = list(range(0, len(ids), max(1, len(ids) // max(1, args.plain_sample))))[:args.plain_sample]
    p_wall = p_sim = 0.0
    for k in idx:
        n, ns, sd, sc = specs[k]
        t1 = time.perf_counter()
        with g.Network(n, ns, sd, thal_scale=sc) as net:
            net.run(steps[k], alg="AB", accum=args.accum, raster=False)
        p_wall += time.perf_counter() - t1
        p_sim += steps[k] / 1000.0
    agg = sim_s / (last - first)
    return {
        "metric": "served: p50/p99 request completion latency and aggregate simulated seconds per second",
        "value": agg, "unit": "sim-s/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Philox networks, Poisson arrivals)",
        "config": {"workload": f"open-loop Poisson {args.rate} req/s for {args.serve} s, N log-uniform [1e3, 5e4], "
                               f"Nsyn in {{100,500,1000}}, thalamic x U(0.8,1.5), steps U[100,5000]",
                   "accum": args.accum, "chunk": args.chunk, "budget_gb": args.budget_gb,
                   "requests": len(ids), "offered_sim_s_per_s": sim_s / args.serve, "seed": args.seed},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "mean": float(lat.mean()), "max": float(lat.max())},
        "events_per_s": events / (last - first), "wall_s": wall,
        "server": {k: st[k] for k in ("completed", "failed", "chunks", "max_active", "max_bytes", "busy_ms",
                                      "device_ms")},
        "plain": {"sim_s_per_s": p_sim / p_wall, "sample": len(idx), "scheme": "AB",
                  "note": "each request alone: create (device connectivity build) + run, back to back"},
        "speedup_vs_plain": agg / (p_sim / p_wall),
    }


def extras(g, args):
    """Batched and served summaries beside the single-network line, in the
    same accumulation mode: 256 networks in one batch (1000 steps), then
    open-loop serving at 50 req/s for 120 s and an overload run at 1000 req/s
    for 10 s (offered about 2500 simulated seconds per second, above what one
    device serves)."""
    import copy
    t0 = time.perf_counter()
    a = copy.copy(args)
    a.batch, a.steps, a.warmup, a.plain_sample, a.slice_len, a.tune = 256, 1000, 100, 8, 0, True
    b = bench_batch(g, a)
    print(f"batched done in {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    s1 = copy.copy(a)
    s1.serve, s1.rate = 120.0, 50.0
    sv = bench_serve(g, s1)
    s2 = copy.copy(a)
    s2.serve, s2.rate = 10.0, 1000.0
    sat = bench_serve(g, s2)
    print(f"served done in {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    return {
        "batched": {"workload": b["config"]["workload"], "accum": a.accum, "sim_s_per_s": b["value"],
                    "events_per_s": b["events_per_s"], "ms_per_step": b["ms_per_step"],
                    "plain_sim_s_per_s": b["plain"]["sim_s_per_s"], "speedup_vs_plain": b["speedup_vs_plain"],
                    "roofline": b["roofline"]},
        "served": {"workload": sv["config"]["workload"], "accum": a.accum, "p50_s": sv["latency_s"]["p50"],
                   "p99_s": sv["latency_s"]["p99"], "sim_s_per_s": sv["value"],
                   "offered_sim_s_per_s": sv["config"]["offered_sim_s_per_s"],
                   "overload": {"rate_req_s": s2.rate, "sim_s_per_s": sat["value"], "p50_s": sat["latency_s"]["p50"],
                                "p99_s": sat["latency_s"]["p99"], "speedup_vs_plain": sat["speedup_vs_plain"]},
                   "plain_sim_s_per_s": sat["plain"]["sim_s_per_s"]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--nsyn", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=20221)
    ap.add_argument("--regime", type=int, default=1)
    # Default plan: what g4s_autotune picks for this bucket on MI355X
    # (SKL atomic, 256 slices, 1024 threads); --alg auto re-tunes in-process.
    ap.add_argument("--alg", default="SKL", help="scheme, or 'auto' for the autotuner's pick")
    ap.add_argument("--accum", default="atomic")
    ap.add_argument("--slices", type=int, default=256)
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--sweep", action="store_true", help="also time every scheme x mode (stderr)")
    ap.add_argument("--cpu-steps", type=int, default=5000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=0, metavar="COUNT",
                    help="batched form instead: COUNT independent networks (N log-uniform in [1e3, 5e4], "
                         "Nsyn in {100, 500, 1000}, thalamic scale U(0.8, 1.5)) in one batch, compared "
                         "with back-to-back plain runs of a sample of them")
    ap.add_argument("--serve", type=float, default=0.0, metavar="SECONDS",
                    help="served form instead: open-loop Poisson arrivals at --rate req/s for SECONDS, each "
                         "a batched-form network with steps U[100, 5000]; p50/p99 latency, aggregate sim-s/s")
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--slice-len", type=int, default=0, help="batched: neurons per slice, 0 = per instance")
    ap.add_argument("--schedule", default="step", choices=["step", "queue"])
    ap.add_argument("--batch-block", type=int, default=0, help="batched: threads per workgroup (256/512/1024)")
    ap.add_argument("--streams", type=int, default=0,
                    help="batched: instance groups stepped on their own streams, 0 = the library default")
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--tune", action="store_true",
                    help="batched: tune the slice length per (N, Nsyn, drive) bucket on the mix first")
    ap.add_argument("--tune-trials", type=int, default=40, help="batched: trials of the mix tuner")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched (256 networks) and served (Poisson arrivals) summaries")
    ap.add_argument("--headline-only", action="store_true",
                    help="time only the headline window (no raster, kernel-event or zero-spike legs, no "
                         "extras, no CPU baseline): the command profiled for the counter summaries")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times with a plain copy kernel "
                         "(reference reading for the FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.copy_calibrate > 0:
        import torch
        nbytes = args.copy_calibrate << 20
        a = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": 5}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (independent networks shard by "
              "placing whole networks on GPUs, with no traffic between them). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import gpu4snn_amd as g

    if args.batch > 0:
        print(json.dumps(bench_batch(g, args)), flush=True)
        return
    if args.serve > 0:
        print(json.dumps(bench_serve(g, args)), flush=True)
        return

    t_build = time.perf_counter()
    net = g.Network(args.n, args.nsyn, args.seed, regime=args.regime)
    t_build = time.perf_counter() - t_build
    print(f"network N={args.n} Nsyn={args.nsyn} built+uploaded in {t_build:.1f} s", file=sys.stderr, flush=True)

    if args.sweep:
        for accum in ("atomic", "det"):
            for alg in ("AP", "N", "S", "AB", "SKL"):
                print(f"sweep {alg}/{accum} ...", file=sys.stderr, flush=True)
                net.reset()
                net.run(args.warmup, alg=alg, accum=accum, raster=False)
                _, c = net.run(args.steps, alg=alg, accum=accum, raster=False)
                tm = net.timing()
                ms = tm["device_ms"] / args.steps
                by = sum(step_bytes(args.n, args.nsyn, int(x)) for x in c)
                print(f"sweep {alg:>3}/{accum:<6} {ms*1000:8.2f} us/step  {ms*1000:8.1f} ms/sim-s  "
                      f"{c.sum()*args.nsyn/(tm['device_ms']/1e3):.3e} ev/s  "
                      f"{by/(tm['device_ms']/1e3)/1e9:8.1f} GB/s  spikes/step {c.mean():.0f}",
                      file=sys.stderr, flush=True)

    plan_info = None
    if args.alg == "auto":
        t_tune = time.perf_counter()
        plan = net.autotune(args.accum, trial_steps=100)
        plan_info = plan.as_dict()
        plan_info["tune_s"] = time.perf_counter() - t_tune
        alg, block, slices = g.ALG_NAMES[plan.alg], plan.opts.block, plan.opts.slices
        print(f"autotune: {plan_info}", file=sys.stderr, flush=True)
    else:
        alg, block, slices = args.alg, args.block, args.slices
    kw = dict(alg=alg, accum=args.accum, block=block, slices=slices)

    def window(raster=False, kernel_events=False):
        """The timed window: from reset, args.warmup untimed steps, then
        args.steps steps; returns (wall seconds, counts, timing).  The same
        window runs once untimed first, so that host-side first-use costs
        (buffers, launch set-up for this step count) stay out of the one
        timed run; the timed steps start from the same reset state."""
        net.reset()
        net.run(args.warmup, raster=raster, **kw)
        net.run(args.steps, raster=raster, kernel_events=kernel_events, **kw)
        net.reset()
        net.run(args.warmup, raster=raster, **kw)  # the raster leg's copies warm up too
        t = time.perf_counter()
        _, c = net.run(args.steps, raster=raster, kernel_events=kernel_events, **kw)
        return time.perf_counter() - t, c, net.timing()

    wall, counts, tm = window()
    spikes = int(counts.sum())
    events = spikes * args.nsyn
    if args.headline_only:
        print(json.dumps({"metric": "synaptic events/s, headline window only", "value": events / wall,
                          "unit": "synaptic events/s", "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": wall * 1e3 / args.steps, "device_ms_per_step": tm["device_ms"] / args.steps,
                          "scheme": alg, "accum": args.accum, "block": block, "slices": slices,
                          "spikes_per_step": spikes / args.steps,
                          "kernel_bytes_per_step": sum(kernel_bytes(alg, args.accum, args.n, args.nsyn, int(x))
                                                       for x in counts) / args.steps,
                          "step_bytes_per_step": sum(step_bytes(args.n, args.nsyn, int(x))
                                                     for x in counts) / args.steps,
                          "launch": net.launch_info()}), flush=True)
        net.close()
        return
    # the same window with the raster rows copied to host memory
    wall_r, _, _ = window(raster=True)
    # Roofline of the dominant kernel: the same window again, each launch of
    # that kernel bracketed by events stamped by its own dispatch packet on
    # the library's stream
    _, kc, tk = window(kernel_events=True)
    kb = sum(kernel_bytes(alg, args.accum, args.n, args.nsyn, int(x)) for x in kc)
    k_launches = max(1, tk["kernel_launches"])
    per_launch_bytes = kb / k_launches
    avg_launch_s = tk["kernel_ms"] / 1e3 / k_launches
    achieved = per_launch_bytes / avg_launch_s / 1e9
    kname = KERNEL_OF.get((alg, args.accum), "step")
    steps_per_launch = args.steps / k_launches
    traffic, traffic_ratio = traffic_for_window(
        profiled_traffic(kname, alg, args.accum, args.n, args.nsyn, args.regime, slices, block),
        kb / args.steps, steps_per_launch)
    # The step's floor: the same window with the thalamic input off (no
    # spikes), i.e. update, bitmap scan and grid barrier without propagation
    net.set_regime(args.regime, 0.0)
    _, c0, t0m = window()
    net.set_regime(args.regime, 1.0)
    floor_us = t0m["device_ms"] * 1e3 / args.steps
    step_us = tm["device_ms"] * 1e3 / args.steps
    linfo = net.launch_info()

    out = {
        "metric": "synaptic events/s + ms per sim-second, N=100k; p50/p99 req latency (served)",
        "value": events / wall,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "ms_per_sim_second": wall * 1e3 / args.steps * 1000.0,
        "device_ms_per_step": tm["device_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox network and thalamic stream)",
        "config": {"workload": f"single network N={args.n} Nsyn={args.nsyn} regime={args.regime} "
                               f"({'quiet balanced irregular'.split()[args.regime]}), 1 ms steps",
                   "scheme": alg, "accum": args.accum, "block": block, "slices": slices,
                   "autotune": plan_info, "spikes_per_step": spikes / args.steps,
                   "firing_rate_hz": spikes / args.steps / args.n * 1000.0, "seed": args.seed},
        "ms_per_step_with_raster": wall_r * 1e3 / args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_ratio": traffic_ratio,
                     "kernel": kname, "bytes_per_launch": per_launch_bytes, "avg_launch_us": avg_launch_s * 1e6,
                     "steps_per_launch": steps_per_launch, "kernel_us_per_step": avg_launch_s * 1e6 / steps_per_launch,
                     "launches_timed": k_launches, "window": f"steps {args.warmup}..{args.warmup + args.steps} from reset",
                     "launch": linfo},
        "step_decomposition": {"device_us_per_step": step_us, "zero_spike_us_per_step": floor_us,
                               "propagation_us_per_step": step_us - floor_us,
                               "zero_spike_spikes": int(c0.sum()),
                               "note": "zero_spike: the same plan and window with the thalamic input off (update, "
                                       "bitmap scan, grid barrier); 40% of the HBM roofline would need "
                                       f"{kb / args.steps / (0.4 * HBM_PEAK_GBS * 1e9) * 1e6:.2f} us per step"},
    }
    net.close()
    if not args.no_extras:
        out.update(extras(g, args))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n, args.nsyn, args.seed, args.regime, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
