"""The drop-in CLI on the GPU against the published golden row and the oracle.

Runs axon_amd/bin/axon with the reference's own command line and checks
(1) the five published numbers of the Config C seed-0 12-pattern row
(README.md:57-66), extracted the way bench/cap10.sh:14-27 extracts them, and
(2) that every report line the oracle prints for the same run appears
verbatim, in order, in the CLI output."""
import json
import os
import re
import subprocess

import pytest

import axon_amd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PUB = json.load(open(os.path.join(HERE, "golden", "published.json")))
ORA = json.load(open(os.path.join(HERE, "golden", "oracle_reports.json")))


def run_cli(args, timeout=300):
    print("cli:", " ".join(args), flush=True)
    r = subprocess.run([axon_amd.CLI_PATH] + list(args), capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def cap10_fields(out):
    """bench/cap10.sh:14-27, restated with Python string handling."""
    lines = out.splitlines()
    w = re.search(r"within-assembly :.*= ([0-9.]+)%", out).group(1)
    i = [k for k, l in enumerate(lines) if "pattern completion with bias OFF" in l][0]
    row = [l for l in lines[i + 1:i + 5] if l.startswith("   75%")][0].split()
    s, o, m = row[2], row[3], row[4]
    j = [k for k, l in enumerate(lines) if "WAVE:" in l][0]
    free = []
    for l in lines[j + 1:]:  # sed -n '/WAVE:/,/^=====/p': the range end is looked for after the start line
        if l.startswith("====="):
            break
        p = l.split()
        if len(p) > 4 and p[1] == "FREE":
            free.append(p[4])
    sustained = f"{float(free[-1]) / float(free[0]):.3f}"
    # grep -A1 'free-phase drive' | tail -1: the line after the last match
    k = [k for k, l in enumerate(lines) if "free-phase drive" in l][-1]
    d = [t.split(":")[1] for t in lines[k + 1].split() if t.startswith("t6:")][0]
    return w, s, o, m, sustained, d


def assert_lines_in_order(out, expected):
    got = out.splitlines()
    pos = 0
    for line in expected:
        if not line.strip():
            continue
        try:
            pos = got.index(line, pos) + 1
        except ValueError:
            raise AssertionError(f"oracle line not found in order: {line!r}")


def test_golden_row_full_scale():
    g = PUB["golden_row"]
    out = run_cli(ORA["c1_trained"]["args"])
    w, s, o, m, sus, d = cap10_fields(out)
    assert (w, s, o, sus, d) == (g["within_pct"], g["recall"], g["best_other"], g["sustained"], g["drive_t6"])
    assert_lines_in_order(out, ORA["c1_trained"]["report"])


def test_untrained_control_full_scale():
    out = run_cli(ORA["c1_untrained"]["args"])
    w, s, o, m, sus, d = cap10_fields(out)
    assert w == PUB["untrained"]["within_pct"] and d == PUB["untrained"]["drive_t6"]
    assert_lines_in_order(out, ORA["c1_untrained"]["report"])


@pytest.mark.parametrize("name", ["c0_small", "c0_defaults", "c0_t20"])
def test_small_config_report_matches_oracle(name):
    """Every report block (settle, epochs, firing steps, diag, A1, A3, A4, A4b,
    A4c, UNIONCHAR, SHARED vs EXCLUSIVE, scaffold, multiplicity, PARITY, WAVE,
    free-run diagnostic, final table): the oracle's lines, in order."""
    out = run_cli(ORA[name]["args"])
    assert len(ORA[name]["report"]) > 150
    assert_lines_in_order(out, ORA[name]["report"])


def test_cli_self_checks():
    out = run_cli(["ncol=512", "incols=128", "stimcols=64", "epochs=1", "vbs=1", "vtopk=1", "quiet=1"])
    assert out.count("EXACT") == 5 and "FAIL" not in out
    assert "MISMATCHES: 0  -> BIT-IDENTICAL" in out


def test_cli_rejects_undefined_config():
    r = subprocess.run([axon_amd.CLI_PATH, "incols=4096"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "incols" in r.stderr
