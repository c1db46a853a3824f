"""Batched sweep driver: the reference's capacity and seed sweeps
(bench/cap10.sh, bench/seeds.sh) with every network trained in shared
launches instead of one CLI run per network.

Each (stored-pattern load, seed) network is created with the sweep's keys,
all of them are trained together (axon_multi_train: epochs x patterns
learning presentations each), and every network is saved as a snapshot.  The
CLI then runs on each snapshot (`load=`): it skips the epochs and runs the rest
of the reference protocol unchanged (the extra presentation, references and
readouts), printing the report lines the reference's scripts parse.  This
module parses those lines as the scripts do and prints their tables.

    python -m axon_amd.sweep --table seeds --stim 24,36,42,48,60 --seeds 0,1,2,3,4 --out /tmp/sweep
    python -m axon_amd.sweep --table cap10 --stim 12,24,36,48,60 --out /tmp/sweep
"""
import argparse
import os
import re
import statistics
import subprocess
import sys
import time

from . import CONFIG_C, Config, MultiNet, Net

HERE = os.path.dirname(os.path.abspath(__file__))
AXON_BIN = os.path.join(HERE, "bin", "axon")

READOUT = "wave=1 wavet=35 deep=1".split()


def wave_free(out: str):
    """First and last FREE row's 5th field inside the WAVE: block (seeds.sh:12-13)."""
    m = re.search(r"WAVE:.*?^=====", out, re.S | re.M)
    rows = [ln.split() for ln in (m.group(0) if m else "").splitlines()]
    vals = [float(r[4]) for r in rows if len(r) > 4 and r[1] == "FREE"]
    return (vals[0], vals[-1]) if vals else (float("nan"), float("nan"))


def completion75(out: str):
    """(self, other, margin) of the 75% row after 'pattern completion with bias OFF' (cap10.sh:17-19)."""
    lines = out.splitlines()
    for i, ln in enumerate(lines):
        if "pattern completion with bias OFF" in ln:
            for r in lines[i + 1:i + 5]:
                if r.startswith("   75%"):
                    f = r.split()
                    return float(f[2]), float(f[3]), float(f[4])
    return float("nan"), float("nan"), float("nan")


def within(out: str) -> float:
    """Density after 'within-assembly :' (cap10.sh:16)."""
    for ln in out.splitlines():
        if "within-assembly :" in ln:
            m = re.search(r"= ([0-9.]+)%", ln)
            if m:
                return float(m.group(1))
    return float("nan")


def drive6(out: str) -> float:
    """t6 of the line after 'free-phase drive' (cap10.sh:22)."""
    lines = out.splitlines()
    for i, ln in enumerate(lines):
        if "free-phase drive" in ln and i + 1 < len(lines):
            for tok in lines[i + 1].split():
                if tok.startswith("t6:"):
                    return float(tok[3:])
    return float("nan")


def jaccard(out: str) -> float:
    """'measured mean pairwise Jaccard = x' (cap10.sh:23)."""
    m = re.search(r"measured mean pairwise Jaccard.*?= ([0-9.]+)", out)
    return float(m.group(1)) if m else float("nan")


def sum_union(out: str) -> str:
    """'sum of assembly sizes S | union U | excess E ->' as 'S|U|E' (cap10.sh:24)."""
    for ln in out.splitlines():
        if "sum of assembly sizes" in ln:
            s = re.sub(r".*sizes ", "", ln)
            s = re.sub(r" -> .*", "", s)
            return s.replace(" | ", "|").replace("union ", "").replace("excess ", "")
    return ""


def train_and_save(keys, specs, epochs, out_dir, log):
    """Create every (stim, seed) network, train them in shared launches, save
    each after its epochs; returns (snapshot paths, training seconds)."""
    cfgs = [Config.parse(keys + [f"stim={n}", f"epochs={epochs}", f"seed={sd}"] + READOUT) for n, sd in specs]
    nets = [Net(c) for c in cfgs]
    mb = MultiNet(nets)
    t0 = time.perf_counter()
    mb.train([epochs] * len(nets), extra=False)
    secs = time.perf_counter() - t0
    mb.close()
    paths = []
    for net, (n, sd) in zip(nets, specs):
        p = os.path.join(out_dir, f"stim{n}_seed{sd}.axs")
        net.save(p)
        net.close()
        paths.append(p)
    log(f"trained {len(specs)} networks together in {secs:.2f} s; snapshots in {out_dir}")
    return paths, secs


def run_cli(keys, n, sd, epochs, snap, out_dir):
    args = [AXON_BIN] + keys + [f"stim={n}", f"epochs={epochs}", f"seed={sd}"] + READOUT + [f"load={snap}"]
    r = subprocess.run(args, capture_output=True, text=True, check=False)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(args)} exited {r.returncode}: {r.stderr.strip()}")
    with open(os.path.join(out_dir, f"stim{n}_seed{sd}.txt"), "w") as f:
        f.write(r.stdout)
    return r.stdout


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--table", choices=("seeds", "cap10"), default="seeds")
    ap.add_argument("--stim", default="24,36,42,48,60", help="comma list of stored-pattern loads")
    ap.add_argument("--seeds", default="0,1,2,3,4")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--keys", default="", help="extra key=value arguments after Configuration C")
    ap.add_argument("--out", required=True, help="directory for snapshots and per-network reports")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    keys = CONFIG_C + a.keys.split()
    loads = [int(x) for x in a.stim.split(",")]
    seeds = [int(x) for x in a.seeds.split(",")] if a.table == "seeds" else [0]
    specs = [(n, sd) for n in loads for sd in seeds]
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    paths, _ = train_and_save(keys, specs, a.epochs, a.out, log)
    outs = {}
    for (n, sd), p in zip(specs, paths):
        outs[(n, sd)] = run_cli(keys, n, sd, a.epochs, p, a.out)
        log(f"  stim={n} seed={sd}: report {os.path.join(a.out, f'stim{n}_seed{sd}.txt')}")
    if a.table == "seeds":  # bench/seeds.sh:7-27
        print(f"=== SUSTAIN VARIANCE: {len(seeds)} seeds x NSTIM {{{','.join(map(str, loads))}}} ===")
        for n in loads:
            v, m, d = [], [], []
            for sd in seeds:
                o = outs[(n, sd)]
                first, last = wave_free(o)
                v.append(round(last / first, 3))
                m.append(completion75(o)[2])
                d.append(within(o))
            sdv = statistics.stdev(v) if len(v) > 1 else 0.0
            sdm = statistics.stdev(m) if len(m) > 1 else 0.0
            sdd = statistics.stdev(d) if len(d) > 1 else 0.0
            print(f"  NSTIM={n}  sustain mean {statistics.mean(v):.3f} sd {sdv:.3f} range [{min(v):.3f},{max(v):.3f}]"
                  f"  |  margin mean {statistics.mean(m):+.3f} sd {sdm:.3f}  |  within% mean {statistics.mean(d):.2f}"
                  f" sd {sdd:.2f}")
            print(f"     raw sustain: {v}")
    else:  # bench/cap10.sh trained rows (the untrained rows need no training to share)
        print("=== CAPACITY SWEEP, configuration C, networks trained together ===")
        print(f"{'stim':<6} {'within%':>7} {'r75self':>8} {'r75oth':>8} {'margin':>8} {'sustain':>9} {'drive6':>8} "
              f"{'asmJacc':>8} {'precond':<9} sum|union|excess")
        for n in loads:
            o = outs[(n, 0)]
            s, oth, mg = completion75(o)
            first, last = wave_free(o)
            d6 = drive6(o)
            pc = "VALID" if (d6 > 58 and s > oth) else ("FAIL:drv" if d6 <= 58 else "FAIL:s<=o")
            print(f"{n:<6} {within(o):>7} {s:>8} {oth:>8} {mg:>8} {last / first:>9.3f} {d6:>8} {jaccard(o):>8} "
                  f"{pc:<9} {sum_union(o)}")


if __name__ == "__main__":
    main()
