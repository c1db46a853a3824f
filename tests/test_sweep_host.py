"""Host side of the sweep driver, on the oracle's reports (no GPU): the report
parsers read every field the reference's table scripts read, and the table
formatters aggregate them as those scripts do (bench/seeds.sh,
bench/table4_gaintraj.sh, bench/table5_union.sh)."""
import json
import math
import os
import statistics

from axon_amd import sweep

HERE = os.path.dirname(os.path.abspath(__file__))
ORA = json.load(open(os.path.join(HERE, "golden", "oracle_reports.json")))


def report(name):
    return "\n".join(ORA[name]["report"]) + "\n"


def test_parsers_read_every_field_of_the_full_report():
    out = report("c1_trained")
    pin = ORA["c1_trained"]["pinned"]
    assert sweep.within(out) == float(pin["within"])
    s, o, m = sweep.completion75(out)
    assert (s, o) == (float(pin["recall"]), float(pin["other"]))
    first, last = sweep.wave_free(out)
    assert f"{last / first:.3f}" == pin["sustained"]
    assert sweep.drive6(out) == float(pin["drive6"])  # the deep block's trace, after the split block's lines
    assert not math.isnan(sweep.jaccard(out)) and sweep.sum_union(out).count("|") == 2
    assert len(sweep.split_lines(out)) == 3
    assert all(not math.isnan(x) for x in sweep.split_fields(out))
    assert len(sweep.exclusive_drive(out)) == 30  # free steps 5..34
    assert len(sweep.unionchar(out)) == 7


def test_table_formatters_aggregate_like_the_scripts():
    a, b = report("c1_trained"), report("c1_untrained")
    outs = {(12, 0, 40): a, (12, 1, 40): b}
    lines, rows = sweep.format_table("seeds", [12], [0, 1], 40, outs)
    v = []
    for o in (a, b):
        f, l_ = sweep.wave_free(o)
        v.append(round(l_ / f, 3))
    assert rows[12]["sustain_mean"] == statistics.mean(v) and rows[12]["sustain_sd"] == statistics.stdev(v)
    assert lines[1].strip().startswith("NSTIM=12")
    lines, rows = sweep.format_table("table4", [12], [0, 1], 40, outs)
    d = [sweep.exclusive_drive(o) for o in (a, b)]
    want = statistics.mean(x[3] / x[1] for x in d if x[1] > 0)  # t6 -> t8
    assert rows[12][0] == want
    outs5 = {(12, 0, 40): a, (12, 1, 40): a, (36, 0, 0): b, (36, 1, 0): b, (36, 2, 0): b}
    lines, rows = sweep.format_table("table5", [12], [0, 1], 40, outs5)
    u = sweep.unionchar(a)
    assert rows[(12, 40)][:7] == u and abs(rows[(12, 40)][7] - 100 * u[1] / (u[0] + u[1])) < 1e-12
    assert rows[(36, 0)][:7] == sweep.unionchar(b)
    assert any(ln.startswith("  UNTRAINED patterns=36") for ln in lines)
