"""Batched sweep driver: the reference's table scripts (bench/cap10.sh,
bench/seeds.sh, bench/table3_subgroups.sh, bench/table4_gaintraj.sh,
bench/table5_union.sh) with every network of a table trained together in
shared launches instead of one CLI run per network.

Each (stored-pattern load, seed) network of a table is created with the
table's keys and all of them are trained together (axon_multi_train: epochs
x patterns learning presentations each).  The rest of the reference protocol
-- the extra learning presentation, the references and the readouts the
table needs -- then runs in-process on each trained network (axon_run with
that network), and its report is parsed the way the script parses it.  The
tables are printed in the scripts' formats.

    python -m axon_amd.sweep --table seeds  --out /tmp/sweep   # Table 2
    python -m axon_amd.sweep --table cap10  --out /tmp/sweep   # Table 1
    python -m axon_amd.sweep --table table3 --out /tmp/sweep
    python -m axon_amd.sweep --table table4 --out /tmp/sweep
    python -m axon_amd.sweep --table table5 --out /tmp/sweep

--stim / --seeds / --epochs / --keys override a table's loads, seeds, epochs
and add keys (e.g. "ncol=512 incols=128" for a small network).
"""
import argparse
import os
import re
import statistics
import sys
import time

from . import CONFIG_C, Config, MultiNet, Net, run

HERE = os.path.dirname(os.path.abspath(__file__))
AXON_BIN = os.path.join(HERE, "bin", "axon")

READOUT = "wave=1 wavet=35 deep=1".split()

# table -> (default loads, default seeds, readout keys, untrained control rows (load, seeds))
TABLES = {
    "cap10": ((12, 24, 36, 48, 60, 74, 90, 140, 200, 300), (0,), READOUT, ()),           # bench/cap10.sh
    "seeds": ((24, 36, 42, 48, 60), (0, 1, 2, 3, 4), READOUT, ()),                        # bench/seeds.sh
    "table3": ((12, 24, 36, 48, 60), (0,), ["split=1"], ((48, (0,)),)),                  # table3_subgroups.sh
    "table4": ((12, 36, 60), (0, 1, 2, 3, 4), ["split=1", "wave=1", "wavet=35"], ()),     # table4_gaintraj.sh
    "table5": ((12, 36, 60), (0, 1, 2, 3, 4), ["unionchar=1"], ((36, (0, 1, 2)),)),       # table5_union.sh
}


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
    """t6 of the line after the last 'free-phase drive' (cap10.sh:22: grep -A1 | tail -1)."""
    lines = out.splitlines()
    hits = [i for i, ln in enumerate(lines) if "free-phase drive" in ln]
    if hits and hits[-1] + 1 < len(lines):
        for tok in lines[hits[-1] + 1].split():
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


def split_lines(out: str):
    """Lines 2-4 of the SHARED vs EXCLUSIVE block (table3_subgroups.sh:14)."""
    m = re.search(r"^[^\n]*SHARED vs EXCLUSIVE[^\n]*\n(.*?)^=====", out, re.S | re.M)
    return (m.group(1) if m else "").splitlines()[:3]


def split_fields(out: str):
    """(shared fraction, shared density %, exclusive density %) of the SHARED vs EXCLUSIVE block."""
    f = re.search(r"shared fraction of assembly : ([0-9.]+)", out)
    s = re.search(r"within-assembly density  SHARED +([0-9.]+)%", out)
    e = re.search(r"within-assembly density  EXCLUSIVE ([0-9.]+)%", out)
    return tuple(float(x.group(1)) if x else float("nan") for x in (f, s, e))


def exclusive_drive(out: str):
    """The values after 'free-phase drive  EXCLUSIVE:' (table4_gaintraj.sh:18)."""
    m = re.search(r"free-phase drive  EXCLUSIVE:(.*)$", out, re.M)
    return [float(x) for x in m.group(1).split()] if m else []


def unionchar(out: str):
    """The seven fields of the UNIONCHAR line (table5_union.sh:19)."""
    m = re.search(r"^UNIONCHAR (.*)$", out, re.M)
    return [float(x) for x in m.group(1).split()] if m else [float("nan")] * 7


def train_together(keys, specs, log):
    """Create every (load, seed, epochs) network and train the ones with
    epochs together in shared launches (no extra presentation: the report's
    protocol runs it).  Returns the networks and the training seconds."""
    nets = [Net(Config.parse(keys + [f"stim={n}", f"epochs={ep}", f"seed={sd}"])) for n, sd, ep in specs]
    todo = [(net, ep) for net, (_, _, ep) in zip(nets, specs) if ep > 0]
    t0 = time.perf_counter()
    if todo:
        mb = MultiNet([net for net, _ in todo])
        mb.train([ep for _, ep in todo], extra=False)
        mb.close()
    secs = time.perf_counter() - t0
    log(f"trained {len(todo)} networks together in {secs:.2f} s")
    return nets, secs


def report(keys, spec, readout, net, out_dir):
    """The rest of the protocol and its report on one trained network (axon_run)."""
    n, sd, ep = spec
    path = os.path.join(out_dir, f"stim{n}_seed{sd}_ep{ep}.txt")
    run(keys + [f"stim={n}", f"epochs={ep}", f"seed={sd}"] + readout, trained=net, report_path=path)
    with open(path) as f:
        return f.read()


def run_table(table, keys, loads, seeds, epochs, out_dir, log):
    """Train a table's networks together, report each, return {(load, seed, epochs): report}."""
    _, _, readout, controls = TABLES[table]
    specs = [(n, sd, epochs) for n in loads for sd in seeds]
    specs += [(n, sd, 0) for n, sds in controls for sd in sds]
    nets, _ = train_together(keys, specs, log)
    outs = {}
    for spec, net in zip(specs, nets):
        outs[spec] = report(keys, spec, readout, net, out_dir)
        net.close()
    return outs


def format_table(table, loads, seeds, epochs, outs):
    """The table as the reference script prints it; returns (lines, parsed rows)."""
    lines, rows = [], {}
    if table == "seeds":  # bench/seeds.sh:7-27
        lines.append(f"=== SUSTAIN VARIANCE: {len(seeds)} seeds x NSTIM {{{','.join(map(str, loads))}}} ===")
        for n in loads:
            v, m, d = [], [], []
            for sd in seeds:
                o = outs[(n, sd, epochs)]
                first, last = wave_free(o)
                v.append(round(last / first, 3))
                m.append(completion75(o)[2])
                d.append(within(o))
            sdv = statistics.stdev(v) if len(v) > 1 else 0.0
            sdm = statistics.stdev(m) if len(m) > 1 else 0.0
            sdd = statistics.stdev(d) if len(d) > 1 else 0.0
            rows[n] = dict(sustain_mean=statistics.mean(v), sustain_sd=sdv, margin_mean=statistics.mean(m),
                           margin_sd=sdm, within_mean=statistics.mean(d), within_sd=sdd, raw=v)
            lines.append(f"  NSTIM={n}  sustain mean {statistics.mean(v):.3f} sd {sdv:.3f} range [{min(v):.3f},"
                         f"{max(v):.3f}]  |  margin mean {statistics.mean(m):+.3f} sd {sdm:.3f}  |  within% mean "
                         f"{statistics.mean(d):.2f} sd {sdd:.2f}")
            lines.append(f"     raw sustain: {v}")
    elif table == "cap10":  # bench/cap10.sh trained rows
        lines.append("=== CAPACITY SWEEP, configuration C, networks trained together ===")
        lines.append(f"{'stim':<6} {'within%':>7} {'r75self':>8} {'r75oth':>8} {'margin':>8} {'sustain':>9} "
                     f"{'drive6':>8} {'asmJacc':>8} {'precond':<9} sum|union|excess")
        for n in loads:
            o = outs[(n, seeds[0], epochs)]
            s, oth, mg = completion75(o)
            first, last = wave_free(o)
            d6 = drive6(o)
            pc = "VALID" if (d6 > 58 and s > oth) else ("FAIL:drv" if d6 <= 58 else "FAIL:s<=o")
            rows[n] = dict(within=within(o), recall=s, other=oth, margin=mg, sustained=round(last / first, 3),
                           drive=d6)
            lines.append(f"{n:<6} {within(o):>7} {s:>8} {oth:>8} {mg:>8} {last / first:>9.3f} {d6:>8} "
                         f"{jaccard(o):>8} {pc:<9} {sum_union(o)}")
    elif table == "table3":  # bench/table3_subgroups.sh
        lines.append("=== Table 3: shared vs exclusive members ===")
        for n in loads:
            lines.append(f"--- patterns={n} ---")
            lines.extend(split_lines(outs[(n, seeds[0], epochs)]))
            rows[(n, epochs)] = split_fields(outs[(n, seeds[0], epochs)])
        for n, sds in TABLES["table3"][3]:
            lines.append(f"--- untrained control, patterns={n} ---")
            lines.extend(split_lines(outs[(n, sds[0], 0)]))
            rows[(n, 0)] = split_fields(outs[(n, sds[0], 0)])
    elif table == "table4":  # bench/table4_gaintraj.sh
        lines.append("=== Table 4: loop gain per two-step window ===")
        lines.append(f"{'patterns':>9} " + " ".join(f"t{t}->{t + 2}" for t in range(6, 28, 4)))
        for n in loads:
            drv = [exclusive_drive(outs[(n, sd, epochs)]) for sd in seeds]
            gains = []
            for t in range(6, 28, 4):
                g = [r[t - 5 + 2] / r[t - 5] for r in drv if t - 5 + 2 < len(r) and r[t - 5] > 0]
                gains.append(statistics.mean(g) if g else float("nan"))
            rows[n] = gains
            lines.append(f"{n:>9} " + "   ".join(f"{g:.3f}" for g in gains))
    elif table == "table5":  # bench/table5_union.sh
        lines.append("=== Table 5: visited-population characterisation ===")
        lines.append(f"{'condition':<22} {'in-asm':>8} {'outside':>8} {'%outside':>9} {'dens(in)':>10} "
                     f"{'dens(OUT)':>11} {'dens(all)':>10} {'mult(OUT)/pop':>14}")

        def agg(reports):
            f = [unionchar(o) for o in reports]
            m = [statistics.mean(x[i] for x in f) for i in range(7)]
            return m + [100 * m[1] / (m[0] + m[1])]

        for n in loads:
            m = agg([outs[(n, sd, epochs)] for sd in seeds])
            rows[(n, epochs)] = m
            lines.append(f"  trained patterns={n:<3} {m[0]:8.0f} {m[1]:8.0f} {m[7]:8.1f}% {m[2]:9.2f}% "
                         f"{m[3]:10.2f}% {m[4]:9.2f}% {m[5]:7.2f}/{m[6]:.2f}")
        for n, sds in TABLES["table5"][3]:
            m = agg([outs[(n, sd, 0)] for sd in sds])
            rows[(n, 0)] = m
            lines.append(f"  UNTRAINED patterns={n}  {m[0]:8.0f} {m[1]:8.0f} {m[7]:8.1f}% {m[2]:9.2f}% "
                         f"{m[3]:10.2f}% {m[4]:9.2f}% {m[5]:7.2f}/{m[6]:.2f}")
    return lines, rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--table", choices=sorted(TABLES), default="seeds")
    ap.add_argument("--stim", default="", help="comma list of stored-pattern loads (default: the table's)")
    ap.add_argument("--seeds", default="", help="comma list of seeds (default: the table's)")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--keys", default="", help="extra key=value arguments after Configuration C")
    ap.add_argument("--out", required=True, help="directory for the per-network reports")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    keys = CONFIG_C + a.keys.split()
    loads = [int(x) for x in a.stim.split(",")] if a.stim else list(TABLES[a.table][0])
    seeds = [int(x) for x in a.seeds.split(",")] if a.seeds else list(TABLES[a.table][1])
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    outs = run_table(a.table, keys, loads, seeds, a.epochs, a.out, log)
    lines, _ = format_table(a.table, loads, seeds, a.epochs, outs)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
