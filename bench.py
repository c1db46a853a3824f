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
kernel (per-kernel HIP-event timing on an identical twin network), the
whole-run HBM rate (every kernel's algorithmic bytes over the wall time; also
for the plain and batched many-network runs), and the CPU oracle's
throughput on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "synaptic events/s (GSyn/s) & % HBM peak, batched vs plain; p99 recall latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of this workload's kernels, measured with rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE (means over all dispatches of the
# Configuration C training run), each scaled by what the counter reads for a
# 256 MiB device copy (x2.000 and x1.000); profiles/k_drive_c1_traffic.md
# (taken with 4-column propagation and 16-column emit workgroups, the same
# accesses per column).
# propagation: the average launch, 6 with the previous timestep's plasticity
# (9492173 B), 2 without (7940915 B) and 1 with the first timestep's work
# (5001830 B) per presentation; last_emit: the last timestep's emit with its
# plasticity.
MEASURED_TRAFFIC = {"propagation": 8648522, "topk_emit": 1526784, "last_emit": 6796390}
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


def acc_bytes(cfg):
    """Bytes per accumulator: 2 when none can reach 2^16 (32 branches of at
    most 256 + boost in mode 1, 256 in mode 0, 1 in mode 2, + facil), else 4."""
    per_branch = {0: 256, 1: 256 + cfg.boost}.get(cfg.mode, 1)
    return 2 if 32 * per_branch + cfg.facil <= 0xFFFF else 4


def kernel_bytes(ncol, steps, prop_lines, learn_lines, bias=True, aw=2, firsts=0, last_lines=None, last_emits=0,
                 first_launches=None):
    """Algorithmic HBM bytes per kernel over `steps` learning timesteps
    (bias=False: readout steps without the intrinsic bias, which neither read
    nor update it; aw: bytes per accumulator, acc_bytes; firsts: how many of
    the steps are the first of a presentation without side statistics, which
    read the stimulus (4 B/column) and write the spike, history, margin and
    spiking flag (16 B/column) and the bias read-modify-write (8 B/neuron),
    with no propagation: a launch of their own for first_launches of them
    (None: all), the others' work being done by the next step's propagation).

    The synapse planes are word-major tiles, so a kernel touches only the
    128-byte lines of presynaptic columns that spiked (device line counters).
    propagation: 128 B per W3 line read + conn (128 B/column) + bias read
    (4 B/neuron, learning uses the bias) + acc (aw B/neuron) and bbits
    (4 B/neuron) writes + drive (4 B/column) + fired-column masks (32 B/column);
    plus the previous timestep's plasticity, which runs inside it.
    topk_emit: every workgroup's pass over the column drives for the top-K
    (L2-resident: counted once, 4 B/column) + acc read (aw B/neuron) + bias
    read-modify-write (8 B/neuron) + drive read, spike, history and margin
    writes (16 B/column).
    Plasticity: 1 KiB per rewritten word line (4 planes x 128 B, read +
    written).  last_emit: the last timestep of each presentation, whose emit
    launch also runs that timestep's plasticity (last_emits of the steps;
    last_lines of the word lines, axon_last_step_counters; None: all of
    them), the other timesteps' lines being the propagation's."""
    n = ncol * 32
    bb = 1 if bias else 0
    full = steps - firsts
    last = learn_lines if last_lines is None else last_lines
    prop = 128 * prop_lines + full * (128 * ncol + 4 * bb * n + (aw + 4) * n + 4 * ncol + 32 * ncol)
    prop += 1024 * (learn_lines - last)
    per_emit = aw * n + 8 * bb * n + 20 * ncol
    emit = (full - last_emits) * per_emit
    nf = firsts if first_launches is None else first_launches
    first = nf * (8 * bb * n + 20 * ncol)
    prop += (firsts - nf) * (8 * bb * n + 20 * ncol)
    last_emit = 1024 * last + last_emits * per_emit
    return {"propagation": prop, "first_emit": first, "topk_emit": emit, "last_emit": last_emit}


def whole_run_hbm(ncol, steps, lines, seconds, bias=True, aw=2, firsts=0, last_lines=None, last_emits=0,
                  first_launches=None):
    """Algorithmic HBM bytes of every kernel of a run (kernel_bytes over all its
    timesteps and device line counters) over the run's wall time: the average
    rate of the whole job, idle launches and gaps included."""
    by = kernel_bytes(ncol, steps, lines[0], lines[1], bias, aw, firsts, last_lines, last_emits, first_launches)
    tot = sum(by.values())
    gbps = tot / seconds / 1e9
    return {"bytes": int(tot), "by_kernel": {k: int(v) for k, v in by.items()}, "GBps": round(gbps, 1),
            "peak": HBM_PEAK_GBS, "frac": round(gbps / HBM_PEAK_GBS, 4)}


def probe_batch_hbm(ncol, probes, T, union_lines, acc16, seconds):
    """Algorithmic bytes of a probe batch run: every W3 line the batch needs
    read once per step (union over probes) + conn (128 B/column) once per
    step; per probe and step, propagation's spike-word and activity reads
    (4 + 1/8 B/column), accumulator write (2 or 4 B/neuron) and drive write
    (4 B/column), and the select's accumulator and drive reads and spike,
    activity and history writes (2 or 4 B/neuron + 12 + 1/8 B/column)."""
    n = ncol * 32
    aw = 2 if acc16 else 4
    shared = 128 * union_lines + T * 128 * ncol
    per_probe = probes * T * (2 * aw * n + 20 * ncol + ncol // 4)
    tot = shared + per_probe
    gbps = tot / seconds / 1e9
    return {"bytes": int(tot), "w3_union_bytes": 128 * union_lines, "per_probe_state_bytes": int(per_probe),
            "GBps": round(gbps, 1), "peak": HBM_PEAK_GBS, "frac": round(gbps / HBM_PEAK_GBS, 4)}


def propagation_probes():
    """Counterparts of the reference's probe05 (binary propagation, dense and
    with a fraction of the presynaptic columns active) and probe06 (the same
    number of spikes scattered over neurons or gathered into whole columns),
    run on the engine's own propagation kernel (axon_prop_probe) with
    synthetic activity.  Synapse-ops count 1,024 per 128-byte W3 line read
    (32 presynaptic x 32 postsynaptic neurons); dense = all 8,192 synapses
    of every neuron, the reference's 1.36e12 synapse-ops/s figure.  ncol=4096
    is the reference's geometry (W3 = 128 MiB, which the 256 MiB Infinity
    Cache holds between launches); ncol=131072 puts 4 GiB of W3 behind every
    launch, out of any cache."""
    import axon_amd

    out = {}
    for ncol in (4096, 131072):
        net = axon_amd.Net(axon_amd.Config.parse(CONFIG_C + [f"ncol={ncol}", "incols=512", "stim=2"]))
        rng = np.random.default_rng(ncol)
        n_neur = ncol * 32
        res = {}

        def run(spk, reps=20):
            us, lines = axon_amd.prop_probe(net, spk, reps)
            return {"us": round(us, 3), "lines": lines, "GBps": round(128 * lines / us / 1e3, 1),
                    "gsyn_ops_s": round(1024 * lines / us / 1e3, 1)}

        dense = run(np.full(ncol, 0xFFFFFFFF, np.uint32))
        dense["vs_reference_1.36e12"] = round(dense["gsyn_ops_s"] * 1e9 / 1.36e12, 2)
        res["dense"] = dense
        sweep = {}
        for pm in (10, 30, 50, 100, 250, 500):  # per mille of columns spiking
            spk = np.zeros(ncol, np.uint32)
            cols = rng.choice(ncol, max(1, ncol * pm // 1000), replace=False)
            spk[cols] = rng.integers(1, 1 << 32, len(cols), dtype=np.uint64).astype(np.uint32)
            sweep[f"{pm / 10:g}%"] = run(spk)
        res["active_columns_sweep"] = sweep
        col = {}
        for pct in (0.25, 1.0, 5.0):  # probe06: the same spike count, scattered vs columnar
            nspk = int(n_neur * pct / 100)
            scat = np.zeros(n_neur, bool)
            scat[rng.choice(n_neur, nspk, replace=False)] = True
            scat_w = np.packbits(scat.reshape(ncol, 32), axis=1, bitorder="little").view(np.uint32).ravel()
            clus_w = np.zeros(ncol, np.uint32)
            clus_w[rng.choice(ncol, nspk // 32, replace=False)] = 0xFFFFFFFF
            a, b = run(scat_w), run(clus_w)
            col[f"{pct:g}%"] = {"spikes": nspk, "scattered": a, "columnar": b,
                                "speedup": round(a["us"] / b["us"], 2),
                                "line_ratio": round(a["lines"] / max(1, b["lines"]), 2)}
        res["scattered_vs_columnar"] = col
        out[f"ncol_{ncol}"] = res
        net.close()
    return out


def hbm_calibration():
    """Achieved read bandwidth of the engine's access shapes, beyond the 256 MiB
    Infinity Cache (4 GiB) and within it (one plane, 128 MiB)."""
    import axon_amd

    out = {}
    for name, kind in (("stream", 0), ("line_128B_gather", 1), ("tile_1KiB_gather", 2)):
        for label, nbytes in (("4GiB", 4 << 30), ("128MiB", 128 << 20)):
            g = axon_amd.mem_probe(kind, nbytes, 5)
            out[f"{name}_{label}"] = {"GBps": round(g, 1), "frac_of_peak": round(g / HBM_PEAK_GBS, 4)}
    out["propagation_probes"] = propagation_probes()
    out["sync_probe"] = sync_probes()
    return out


def sync_probes():
    """Microseconds per launch of storing workgroups with no cross-workgroup
    synchronisation, with an arrival atomic per workgroup, with an agent-scope
    release fence before it, and with a full agent-scope fence (what a grid
    barrier or a last-workgroup reduction would add to every workgroup), at
    the shapes of one network's propagation (1 x 1024 workgroups) and of a
    5-network batched propagation (5 x 512): the measurement behind running
    each timestep as separate launches rather than one persistent kernel."""
    import axon_amd

    out = {}
    for label, groups, wgs in (("1x1024", 1, 1024), ("5x512", 5, 512)):
        out[label] = {name: round(axon_amd.sync_probe(kind, groups, wgs, 200), 2)
                      for kind, name in ((0, "plain"), (1, "arrival_atomic"), (2, "release_fence_atomic"),
                                         (3, "full_fence_atomic"))}
    return out


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
    lines = net.line_counters()
    last_lines = net.last_step_counters()[1]
    prop, learn = net.work_counters(reset=True)
    net.close()

    # twin network: same seed and presentation sequence, per-kernel HIP events
    log("bench: per-kernel event timing on a twin network")
    twin = axon_amd.Net(cfg)
    train_presentations(twin, 0, args.warmup, T, tc, nstim)
    twin.sync()
    twin.work_counters(reset=True)
    twin.set_profiling(True)
    train_presentations(twin, args.warmup, args.steps, T, tc, nstim)
    ms, launches = twin.kernel_times(reset=True)
    prop2, learn2 = twin.work_counters(reset=True)
    twin.close()
    assert (prop2, learn2) == (prop, learn), "twin network diverged"
    steps = args.steps * T
    # training presentations are asynchronous: each one's first step is one launch (first_emit)
    # each presentation's last timestep learns: its emit launch runs its plasticity too (last_emit); its
    # first timestep runs inside the second's propagation (no first_emit launches) when clamped
    kb = kernel_bytes(ncol, steps, lines[0], lines[1], aw=acc_bytes(cfg), firsts=args.steps, last_lines=last_lines,
                      last_emits=launches[3], first_launches=launches[1])
    names = ["propagation", "first_emit", "topk_emit", "last_emit"]  # the column top-K runs inside emit
    kern = {}
    for i, nm in enumerate(names):
        if launches[i] == 0:
            continue
        t_s = ms[i] / 1e3
        kern[nm] = {"launches": launches[i], "avg_us": 1e3 * ms[i] / max(launches[i], 1),
                    "bytes_per_launch": kb[nm] / max(launches[i], 1),
                    "GBps": kb[nm] / t_s / 1e9 if t_s > 0 else 0.0}
    dom = max((i for i in range(len(names)) if launches[i]), key=lambda i: ms[i])
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
        "whole_run_hbm": whole_run_hbm(ncol, steps, lines, dt, aw=acc_bytes(cfg), firsts=args.steps,
                                       last_lines=last_lines, last_emits=launches[3], first_launches=launches[1]),
        "roofline": {"bound": "hbm", "kernel": dn, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get(dn),
                     "bytes_per_launch": round(kern[dn]["bytes_per_launch"]),
                     "avg_launch_us": round(kern[dn]["avg_us"], 3)},
    }


def train_alone(net):
    c = net.config
    for ep in range(c.epochs):
        for p in range(c.stim):
            net.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p, outputs=False)
    net.present(0, 1.0, c.T, c.trainclamp, True, False, 31337, outputs=False)
    net.sync()


def probe_set(net, axon_amd):
    """60 stored patterns x {0, 10, 25, 50 %} of their 64 cue columns replaced by
    other input columns, plus 16 random 64-column cues (256 probes)."""
    cues = []
    for r in (0, 6, 16, 32):
        for p in range(net.config.stim):
            cues.append(axon_amd.probe_cue(net, p, r, 100000 + 1000 * r + p))
    for i in range(16):
        cues.append(axon_amd.probe_cue(net, -1, 0, 200000 + i))
    return np.stack(cues)


_POPC8 = np.array([bin(i).count("1") for i in range(256)], np.int64)


def probe_lines(conn, ncol, hists):
    """W3 lines a readout propagation needs, from each probe's spike history:
    at step t a (column, branch) tile reads one 128-byte line per presynaptic
    column of its group (conn) that spiked at step t - 1 (none at step 0).
    Returns (lines summed over probes, lines of the union over probes per
    step): the first is what per-probe reads cost, the second the bytes of
    every tile read once per step for the whole batch."""
    ng = ncol // 8
    tiles_per_group = np.bincount(np.asarray(conn, np.int64), minlength=ng)
    per_probe, union = 0, None
    for h in hists:
        nz = (np.asarray(h)[:-1] != 0).reshape(len(h) - 1, ng, 8)
        gm = np.packbits(nz, axis=2, bitorder="little")[..., 0]
        per_probe += int((_POPC8[gm] * tiles_per_group).sum())
        union = gm.copy() if union is None else union | gm
    return per_probe, int((_POPC8[union] * tiles_per_group).sum())


def c3_probes(args):
    import axon_amd

    cfg = axon_amd.Config.parse(CONFIG_C + "stim=60 epochs=40 seed=0".split())
    log("probes: training the load-60 network")
    net = axon_amd.Net(cfg)
    train_alone(net)
    stims = probe_set(net, axon_amd)
    T, clamp = 35, 5
    log(f"probes: plain path, {len(stims)} probes one presentation each")
    net.line_counters(reset=True)
    t0 = time.perf_counter()
    plain = [net.present(0, 1.0, T, clamp, False, True, 0, 0, stim=s).hist for s in stims]
    t_plain = time.perf_counter() - t0
    lines = net.line_counters()
    tiles_read, _ = net.work_counters(reset=True)
    per_probe_lines, union_lines = probe_lines(net.read("CONN"), net.ncol, plain)
    if args.probe_tile:  # a fixed choice (profiled runs), no sweep
        best, waves, sweep = args.probe_tile, args.probe_waves, {}
    else:
        log("probes: autotuning the probe tile and the propagation's occupancy")
        best, waves, sweep = axon_amd.probe_autotune(net, stims, T, clamp, 0)
    pb = axon_amd.Probes(net, len(stims), T, best, waves)
    acc16 = acc_bytes(cfg) == 2
    pb.run(stims, T, clamp, 0, hist=False)  # warm (graph capture)
    _, hist, ms = pb.run(stims, T, clamp, 0, hist=True)
    exact = all(np.array_equal(hist[i], plain[i]) for i in range(len(stims)))
    pb.close()
    net_ncol = net.ncol
    # what the CPU baseline needs to run the same probes on the same trained network
    cpu_case = {"args": CONFIG_C + "stim=60 epochs=40 seed=0".split(), "T": T, "clamp": clamp, "stims": stims,
                "hist": plain, "state": {nm: net.read(nm).copy() for nm in ("W3", "CONN", "BIAS")}}
    net.close()
    syn = SYN_PER_TILE * tiles_read
    return {"probes": len(stims), "T": T, "plain_gsyn_s": round(syn / t_plain / 1e9, 1),
            "batched_gsyn_s": round(syn / (ms / 1e3) / 1e9, 1), "speedup": round(t_plain / (ms / 1e3), 2),
            "plain_ms": round(1e3 * t_plain, 2), "batched_ms": round(ms, 3), "best_tile": best,
            "best_waves_per_cu": waves,
            "tile_sweep_ms": {f"{t}/{w}": round(v, 3) for (t, w), v in sweep.items()},
            "plain_hbm": whole_run_hbm(net_ncol, len(stims) * T, lines, t_plain, bias=False, aw=acc_bytes(cfg)),
            # W3 lines counted once per (tile, line) per step for the whole batch; the per-probe
            # state (spikes, accumulators at 2 bytes when 16-bit, drives) once per probe
            "batched_hbm": probe_batch_hbm(net_ncol, len(stims), T, union_lines, acc16, ms / 1e3),
            "w3_lines": {"per_probe_sum": per_probe_lines, "union_per_step": union_lines,
                         "device_counter_plain": lines[0], "host_equals_device": per_probe_lines == lines[0]},
            "bit_exact_vs_plain": bool(exact), "_cpu_case": cpu_case}


def c2_networks(args):
    import axon_amd

    loads, seeds = (12, 24, 36, 48, 60), range(5)
    specs = [(P, sd) for P in loads for sd in seeds]
    ep = args.net_epochs
    mk = lambda P, sd: axon_amd.Config.parse(CONFIG_C + f"stim={P} epochs={ep} seed={sd}".split())  # noqa: E731
    names = ("W0", "W1", "W2", "W3", "BIAS")
    log(f"networks: plain path, {len(specs)} networks trained one after another")
    ref, t_plain, tiles_plain, lines_plain = {}, 0.0, 0, [0, 0]
    T = mk(*specs[0]).T
    for P, sd in specs:
        net = axon_amd.Net(mk(P, sd))
        t0 = time.perf_counter()
        train_alone(net)
        t_plain += time.perf_counter() - t0
        ref[(P, sd)] = [axon_amd.checksum(net, nm) for nm in names]
        tiles_plain += net.work_counters()[0]
        lines_plain = [a + b for a, b in zip(lines_plain, net.line_counters())]
        net.close()
    log("networks: batched, all networks in shared launches")
    nets = [axon_amd.Net(mk(P, sd)) for P, sd in specs]
    mb = axon_amd.MultiNet(nets)
    ms = mb.train([ep] * len(nets), extra=True)
    exact = all([axon_amd.checksum(n, nm) for nm in names] == ref[spec] for n, spec in zip(nets, specs))
    tiles = sum(n.work_counters()[0] for n in nets)
    learn = sum(n.work_counters()[1] for n in nets)
    lines = [sum(n.line_counters()[i] for n in nets) for i in (0, 1)]
    ncol = nets[0].ncol
    mb.close()
    for n in nets:
        n.close()
    return {"networks": len(specs), "epochs": ep, "presentations": sum(ep * P + 1 for P, _ in specs),
            "plain_s": round(t_plain, 3), "batched_s": round(ms / 1e3, 3), "speedup": round(t_plain / (ms / 1e3), 2),
            "plain_gsyn_s": round(SYN_PER_TILE * tiles_plain / t_plain / 1e9, 1),
            "batched_gsyn_s": round(SYN_PER_TILE * tiles / (ms / 1e3) / 1e9, 1),
            "batched_plasticity_gsyn_s": round(SYN_PER_TILE * learn / (ms / 1e3) / 1e9, 1),
            "plain_hbm": whole_run_hbm(ncol, T * sum(ep * P + 1 for P, _ in specs), lines_plain, t_plain,
                                       aw=acc_bytes(mk(*specs[0])), firsts=sum(ep * P + 1 for P, _ in specs)),
            "batched_hbm": whole_run_hbm(ncol, T * sum(ep * P + 1 for P, _ in specs), lines, ms / 1e3,
                                         aw=acc_bytes(mk(*specs[0])), firsts=sum(ep * P + 1 for P, _ in specs)),
            "bit_exact_vs_plain": bool(exact and tiles == tiles_plain)}


def c4_serving(args):
    import axon_amd

    loads = (12, 36, 60, 100, 140, 170, 200, 200)
    log("serving: training 8 resident networks together")
    nets = [axon_amd.Net(axon_amd.Config.parse(CONFIG_C + f"stim={P} epochs=40 seed={i} maxt=200".split()))
            for i, P in enumerate(loads)]
    mb = axon_amd.MultiNet(nets)
    mb.train([40] * len(nets), extra=True)
    mb.close()
    out = {"networks": len(nets), "loads": list(loads), "T_range": [10, 200], "rates": []}
    if args.serve_tile == 0:
        # the tuning table's entry for (ncol, slots per network, T about the mean query length) is
        # what a server created with tile 0 takes
        log(f"serving: tuning the probe tile for {args.serve_slots} slots")
        k = 3
        cues = np.stack([axon_amd.probe_cue(nets[k], i % loads[k], (0, 6, 16, 32)[i % 4], 400000 + i)
                         for i in range(args.serve_slots)])
        out["tuned"] = axon_amd.probe_autotune_sizes(nets[k], cues, 100, 5, [args.serve_slots])[0]
    for rate in args.serve_rates:
        n_q = max(200, int(rate * args.serve_seconds))
        rs = np.random.RandomState(rate)  # MT19937 stream
        gaps = rs.exponential(1.0 / rate, n_q)
        arrivals = np.cumsum(gaps)
        qs = []
        for i in range(n_q):
            k = int(rs.randint(0, len(nets)))
            if rs.randint(0, 16) == 0:
                cue = axon_amd.probe_cue(nets[k], -1, 0, 300000 + i)
            else:
                cue = axon_amd.probe_cue(nets[k], int(rs.randint(0, loads[k])), (0, 6, 16, 32)[rs.randint(0, 4)],
                                         300000 + i)
            qs.append((k, cue, int(rs.randint(10, 201))))
        srv = axon_amd.Server(nets, slots_per_net=args.serve_slots, max_t=200, tile=args.serve_tile,
                              steps_per_tick=args.serve_steps)
        out["server_tile"], out["server_waves_per_cu"] = srv.config()["tile"], srv.config()["waves_per_cu"]
        log(f"serving: {n_q} queries at {rate}/s")
        start = srv.clock()
        ids, lat, res, nxt, last_done = {}, [], {}, 0, start
        while len(lat) < n_q:
            now = srv.clock()
            while nxt < n_q and start + arrivals[nxt] <= now:
                k, cue, T = qs[nxt]
                ids[srv.submit(k, cue, T, 5, 0, start + arrivals[nxt])] = nxt
                nxt += 1
            if srv.tick():
                for rid in srv.completed():
                    r = srv.result(rid, 200)
                    i = ids.pop(rid)
                    lat.append(r[2])
                    if i % max(1, n_q // 24) == 0:
                        res[i] = r
                last_done = srv.clock()
        exact = True
        for i, r in res.items():
            k, cue, T = qs[i]
            h = nets[k].present(0, 1.0, T, 5, False, True, 0, 0, stim=cue)
            exact &= bool(np.array_equal(r[0], h.spikes) and np.array_equal(r[1], axon_amd.row_digest(h.hist)))
        srv.close()
        lat_ms = np.array(lat) * 1e3
        out["rates"].append({"rate_qps": rate, "queries": n_q, "qps": round(n_q / (last_done - start), 1),
                             "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                             "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                             "checked_bit_exact": len(res), "bit_exact_vs_alone": exact})
    for n in nets:
        n.close()
    return out


def readouts(args):
    """The readout reductions on the trained full-scale network (12 patterns):
    wall time per entry point call (device work plus the copy of its result)
    and the algorithmic bytes it reads: the 28 recurrent tiles of every column
    in W3 (indeg; within with every recurrent neuron selected) or in all four
    planes (untouched), and the branches of assembly-member columns (wdiag)."""
    import axon_amd

    net = axon_amd.Net(axon_amd.Config.parse(C1))
    c = net.config
    train_presentations(net, 0, c.epochs * c.stim, c.T, c.trainclamp, c.stim)
    net.sync()
    refs = np.stack([net.present(p, 1.0, c.T, c.T, False, True, 2000 + p).spikes for p in range(c.stim)])
    refs[:, :c.incols] = 0
    n, ncol = net.n, net.ncol
    memb = np.zeros(n, np.uint32)
    for q in range(c.stim):
        bits = np.unpackbits(refs[q].view(np.uint8), bitorder="little").astype(np.uint32)
        memb |= bits << np.uint32(q)
    allrec = np.zeros(n, np.uint32)
    allrec[c.incols * 32:] = 1
    rec_w3 = ncol * (32 - c.inbr) * 1024
    member_cols = int(np.count_nonzero(memb.reshape(ncol, 32).any(axis=1)))
    calls = {
        "indeg": (lambda: net.indeg(), rec_w3),
        "untouched": (lambda: net.untouched(), 4 * rec_w3),
        "within_all_recurrent": (lambda: net.within(refs[0], allrec), (ncol - c.incols) * (32 - c.inbr) * 1024),
        "within_one_assembly": (lambda: net.within(refs[0]), 0),
        "wdiag_recurrent": (lambda: net.wdiag(memb, refs.ravel(), c.stim, 0), member_cols * (32 - c.inbr) * 4096),
    }
    out = {}
    for name, (fn, nbytes) in calls.items():
        fn()
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            fn()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        out[name] = {"ms_per_call": round(ms, 3), "algorithmic_bytes": nbytes,
                     "GBps_wall": round(nbytes / (ms / 1e3) / 1e9, 1) if nbytes else None}
    net.close()
    return out


def cpu_networks(args, threads):
    """The many-network workload on the CPU: `threads` of the same networks
    (loads 12-60, seeds 0-4, in that order), one oracle network per host
    thread, each running its training presentations in the plain order for
    the bounded time; aggregate GSyn/s over all of them."""
    import threading

    import oracle

    loads = (12, 24, 36, 48, 60)
    specs = [(P, sd) for sd in range(5) for P in loads][:threads]
    nets = [None] * len(specs)
    done = [0] * len(specs)

    def make(i):
        P, sd = specs[i]
        nets[i] = oracle.OracleNet(oracle.parse(CONFIG_C + f"stim={P} epochs=40 seed={sd}".split()), threads=1)

    ts = [threading.Thread(target=make, args=(i,)) for i in range(len(specs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    T, tc = nets[0].cfg.T, nets[0].cfg.trainclamp
    t0 = time.perf_counter()
    deadline = t0 + args.cpu_seconds

    def train(i):
        P = specs[i][0]
        while time.perf_counter() < deadline:
            p = done[i] % P
            nets[i].present(p, 1.0, T, tc, True, 1000 + p)
            done[i] += 1

    ts = [threading.Thread(target=train, args=(i,)) for i in range(len(specs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    tiles = sum(n.work_counters()[0] for n in nets)
    for n in nets:
        n.close()
    return {"value": round(SYN_PER_TILE * tiles / dt / 1e9, 3), "unit": "GSyn/s", "cores": len(specs),
            "kind": "port",
            "sample": f"{len(specs)} networks (loads 12-60, seeds 0-4), one host thread each, the first "
                      f"{sum(done)} learning presentations in all, {dt:.1f} s"}


def cpu_probes(args, threads, case):
    """The probe workload on the CPU: the oracle network given the GPU-trained
    load-60 network's W3, conn and bias, running the same cues one readout
    presentation at a time (column-parallel over `threads` host threads) for
    the bounded time; every probe it finishes is checked against the GPU's
    history of that probe."""
    import oracle

    ora = oracle.OracleNet(oracle.parse(case["args"]), threads=threads)
    for nm, v in case["state"].items():
        ora.buf(nm)[:] = v
    T, clamp, stims = case["T"], case["clamp"], case["stims"]
    ora.work_counters(reset=True)
    done, same, t0 = 0, True, time.perf_counter()
    while done < len(stims) and time.perf_counter() - t0 < args.cpu_seconds:
        ora.present(0, 1.0, T, clamp, False, 0, 0, stim=stims[done])
        same &= bool(np.array_equal(ora.buf("HIST")[: T * ora.ncol].reshape(T, ora.ncol), case["hist"][done]))
        done += 1
    dt = time.perf_counter() - t0
    tiles = ora.work_counters()[0]
    ora.close()
    return {"value": round(SYN_PER_TILE * tiles / dt / 1e9, 3), "unit": "GSyn/s", "cores": threads, "kind": "port",
            "sample": f"the first {done} of the 256 probes (T={T}) on the GPU-trained load-60 network, "
                      f"{dt:.1f} s", "probes_per_s": round(done / dt, 2), "same_history_as_gpu": same}


def cpu_bench(args, cpu_case=None):
    """The oracle (C++ restatement of the reference kernels), column-parallel on
    the host cores, on the first presentations of the same training run; then
    the many-network and probe workloads (cpu_networks, cpu_probes)."""
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
    out = {"value": round(SYN_PER_TILE * prop / dt / 1e9, 3), "unit": "GSyn/s", "cores": threads,
           "kind": "port",
           "sample": f"first {done} learning presentations ({done * T} timesteps) of the same full-scale "
                     f"training run, {dt:.1f} s",
           "timesteps_per_s": round(done * T / dt, 2)}
    if "networks" in args.sections.split(","):
        log(f"cpu baseline: {threads} networks, one thread each")
        out["networks"] = cpu_networks(args, threads)
    if cpu_case is not None:
        log("cpu baseline: probes on the trained load-60 network")
        out["probes"] = cpu_probes(args, threads, cpu_case)
    return out


def copy_calibration(args):
    """Plain device-to-device copies of a known size (256 MiB read + 256 MiB
    written per launch): the yardstick that scales the FETCH_SIZE / WRITE_SIZE
    counters into bytes for the traffic figure."""
    import torch

    nbytes = 256 << 20
    src = torch.randint(0, 2 ** 31 - 1, (nbytes // 4,), dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(json.dumps({"copy_calibration": {"bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                                           "launches": args.steps, "ok": ok}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480, help="timed learning presentations (480 = 40 epochs)")
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-calib", action="store_true", help="skip the bandwidth and propagation probes")
    ap.add_argument("--sections", default="c1,probes,networks,serving",
                    help="comma list of c1, probes, networks, serving, calib, readouts (copy: counter "
                         "calibration only)")
    ap.add_argument("--net-epochs", type=int, default=40)
    ap.add_argument("--probe-tile", type=int, default=0, help="0: the autotuner's choice")
    ap.add_argument("--probe-waves", type=int, default=0, help="waves per CU with --probe-tile")
    ap.add_argument("--serve-rates", type=lambda x: [int(v) for v in x.split(",")], default=[1000, 5000, 20000])
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-slots", type=int, default=128)
    ap.add_argument("--serve-tile", type=int, default=0, help="0: the autotuner's choice")
    ap.add_argument("--serve-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (every network, probe batch "
              f"and serving queue here lives on one MI355X).", file=sys.stderr)
        sys.exit(2)
    if os.environ.get("RANK", "0") != "0":
        return
    sections = set(args.sections.split(","))
    if "copy" in sections:
        copy_calibration(args)
        return
    g = gpu_bench(args) if "c1" in sections else None
    calib = hbm_calibration() if (("c1" in sections and not args.no_calib) or "calib" in sections) else None
    batched = {}
    if "probes" in sections:
        batched["probes"] = c3_probes(args)
    if "networks" in sections:
        batched["networks"] = c2_networks(args)
    serving = c4_serving(args) if "serving" in sections else None
    reads = readouts(args) if "readouts" in sections else None
    cpu_case = batched.get("probes", {}).pop("_cpu_case", None)
    cpu = None if args.no_cpu else cpu_bench(args, cpu_case)
    if g is None:  # a run of the batched sections alone (profiling): no headline figure
        print(json.dumps({"sections": sorted(sections), "batched_vs_plain": batched, "serving": serving,
                          "readouts": reads, "hbm_calibration": calib, "cpu_baseline": cpu}), flush=True)
        return
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
                   "ncol": 4096, "fan_in": 8192, "patterns": 12, "T": 10, "parallelism": "single GPU",
                   "timed_region": "training presentations without the per-step side statistics (stats "
                                   "counters, drive log) that the reference computes every step; the CLI "
                                   "produces them exactly where it prints them, so no reported output changes"},
        "timesteps_per_s": round(g["timesteps_per_s"], 1),
        "plasticity_gsyn_s": round(g["plasticity_gsyn_s"], 2),
        "kernels": {k: {kk: round(vv, 3) if isinstance(vv, float) else vv for kk, vv in v.items()}
                    for k, v in g["kernels"].items()},
        "roofline": g["roofline"],
        "whole_run_hbm": g["whole_run_hbm"],
        "hbm_calibration": calib,
        "batched_vs_plain": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
