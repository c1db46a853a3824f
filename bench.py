#!/usr/bin/env python3
"""Benchmark of the MI355X AXON engine.

Workload (one GPU): the full-scale single network of the paper's golden row --
131,072 neurons x 8,192 fan-in (1.07 G synapses), Configuration C, 12 stored
patterns, seed 0.  A "step" is one learning presentation of the training
protocol (T = 10 timesteps of propagation -> column top-K -> k-WTA emit ->
bias -> plasticity; pattern p = i mod 12 with stimulus seed 1000 + p), i.e.
exactly what the reference's epoch loop runs (src/axon.mm:1009-1010).

value = synaptic events per second (GSyn/s): 8192 synapses for every
(column, branch) tile propagation evaluates, over the timed region.

Also reported: plasticity synapse updates/s, the roofline of the dominant
kernel (per-kernel HIP-event timing on an identical twin network), and the
CPU oracle's throughput on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "synaptic events/s (GSyn/s) & % HBM peak, batched vs plain; p99 recall latency"
HBM_PEAK_GBS = 8000.0
SYN_PER_TILE = 8192
CONFIG_C = ("nand=4 target=64 k=24 dthr=1 mode=1 inbr=4 stimcols=64 contrast=1 margingate=0 "
            "ltdrep=1 ltdevery=2 biasup=73 biasdown=1 T=10 clampsteps=5").split()
C1 = CONFIG_C + "stim=12 epochs=40 seed=0".split()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def train_presentations(net, start, count, T, tc, nstim, outputs=False):
    for i in range(start, start + count):
        p = i % nstim
        net.present(p, 1.0, T, tc, True, False, 1000 + p, outputs=outputs)


def kernel_bytes(ncol, steps, prop_tiles, learn_tiles):
    """Algorithmic HBM bytes per kernel over `steps` learning timesteps.

    propagation: 1 KiB of W3 per tile read + conn (128 B/column) + bias read
    (4 B/neuron, learning uses the bias) + acc and bbits writes (8 B/neuron)
    + drive (4 B/column).
    emit+plasticity: 8 KiB per rewritten tile (4 planes read + written) +
    acc read (4 B/neuron) + bias read-modify-write (8 B/neuron) + drive read,
    spike, history and margin writes (16 B/column)."""
    n = ncol * 32
    prop = 1024 * prop_tiles + steps * (128 * ncol + 4 * n + 8 * n + 4 * ncol)
    emit = 8192 * learn_tiles + steps * (4 * n + 8 * n + 16 * ncol)
    topk = steps * 3 * 4 * ncol
    return {"propagation": prop, "topk": topk, "emit_plasticity": emit}


def gpu_bench(args):
    import axon_amd

    cfg = axon_amd.Config.parse(C1)
    T, tc, nstim, ncol = cfg.T, cfg.trainclamp, cfg.stim, cfg.ncol
    log(f"bench: create network ncol={ncol} (1.07 G synapses)")
    net = axon_amd.Net(cfg)
    train_presentations(net, 0, args.warmup, T, tc, nstim)
    net.sync()
    net.work_counters(reset=True)
    log(f"bench: timing {args.steps} presentations x {T} timesteps")
    t0 = time.perf_counter()
    train_presentations(net, args.warmup, args.steps, T, tc, nstim)
    net.sync()
    dt = time.perf_counter() - t0
    prop, learn = net.work_counters(reset=True)
    net.close()

    # twin network: same seed and presentation sequence, per-kernel HIP events
    log("bench: per-kernel event timing on a twin network")
    twin = axon_amd.Net(cfg)
    train_presentations(twin, 0, args.warmup, T, tc, nstim)
    twin.sync()
    twin.work_counters(reset=True)
    twin.set_profiling(True)
    train_presentations(twin, args.warmup, args.steps, T, tc, nstim, outputs=True)
    ms, launches = twin.kernel_times(reset=True)
    prop2, learn2 = twin.work_counters(reset=True)
    twin.close()
    assert (prop2, learn2) == (prop, learn), "twin network diverged"
    steps = args.steps * T
    kb = kernel_bytes(ncol, steps, prop, learn)
    names = ["propagation", "topk", "emit_plasticity"]
    kern = {}
    for i, nm in enumerate(names):
        t_s = ms[i] / 1e3
        kern[nm] = {"launches": launches[i], "avg_us": 1e3 * ms[i] / max(launches[i], 1),
                    "bytes_per_launch": kb[nm] / max(launches[i], 1),
                    "GBps": kb[nm] / t_s / 1e9 if t_s > 0 else 0.0}
    dom = max(range(3), key=lambda i: ms[i])
    dn = names[dom]
    achieved = kern[dn]["GBps"]
    return {
        "seconds": dt,
        "gsyn_s": SYN_PER_TILE * prop / dt / 1e9,
        "plasticity_gsyn_s": SYN_PER_TILE * learn / dt / 1e9,
        "timesteps_per_s": steps / dt,
        "ms_per_step": 1e3 * dt / args.steps,
        "prop_tiles": prop, "learn_tiles": learn,
        "kernels": kern,
        "roofline": {"bound": "hbm", "kernel": dn, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": round(kern[dn]["bytes_per_launch"]),
                     "avg_launch_us": round(kern[dn]["avg_us"], 3)},
    }


def cpu_bench(args):
    """The oracle (C++ restatement of the reference kernels), column-parallel on
    the host cores, on the first presentations of the same training run."""
    import oracle

    threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    cfg = oracle.parse(C1)
    log(f"cpu baseline: oracle with {threads} threads, init 1.07 G synapses")
    net = oracle.OracleNet(cfg, threads=threads)
    T, tc, nstim = cfg.T, cfg.trainclamp, cfg.stim
    net.work_counters(reset=True)
    done, t0 = 0, time.perf_counter()
    while True:
        p = done % nstim
        net.present(p, 1.0, T, tc, True, 1000 + p)
        done += 1
        if time.perf_counter() - t0 >= args.cpu_seconds:
            break
    dt = time.perf_counter() - t0
    prop, learn = net.work_counters()
    net.close()
    return {"value": round(SYN_PER_TILE * prop / dt / 1e9, 3), "unit": "GSyn/s", "cores": threads,
            "kind": "port",
            "sample": f"first {done} learning presentations ({done * T} timesteps) of the same full-scale "
                      f"training run, {dt:.1f} s",
            "timesteps_per_s": round(done * T / dt, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480, help="timed learning presentations (480 = 40 epochs)")
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (every network, probe batch "
              f"and serving queue here lives on one MI355X).", file=sys.stderr)
        sys.exit(2)
    if os.environ.get("RANK", "0") != "0":
        return
    g = gpu_bench(args)
    cpu = None if args.no_cpu else cpu_bench(args)
    out = {
        "metric": METRIC,
        "value": round(g["gsyn_s"], 2),
        "unit": "GSyn/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(g["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: seeded reference stimuli and hash-initialised weights (no dataset)",
        "config": {"workload": "full-scale single network training (131072 neurons x 8192 fan-in, 1.07 G "
                               "synapses), Configuration C, 12 patterns, seed 0; step = one 10-timestep "
                               "learning presentation",
                   "ncol": 4096, "fan_in": 8192, "patterns": 12, "T": 10, "parallelism": "single GPU"},
        "timesteps_per_s": round(g["timesteps_per_s"], 1),
        "plasticity_gsyn_s": round(g["plasticity_gsyn_s"], 2),
        "kernels": {k: {kk: round(vv, 3) if isinstance(vv, float) else vv for kk, vv in v.items()}
                    for k, v in g["kernels"].items()},
        "roofline": g["roofline"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
