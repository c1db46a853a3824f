"""Batched sweep: a CLI run that starts from a snapshot (load=) prints the same
report as the full run that saved it (save=), apart from the training lines;
networks trained together in shared launches save the same snapshot as the
CLI's own training; the sweep driver's table parses those reports."""
import os
import subprocess

import numpy as np
import pytest

import axon_amd
from axon_amd import sweep

pytestmark = pytest.mark.gpu

SMALL = ["ncol=512", "incols=128"]
KEYS = sweep.CONFIG_C + SMALL


def cli(args):
    r = subprocess.run([sweep.AXON_BIN] + args, capture_output=True, text=True, check=False)
    assert r.returncode == 0, r.stderr
    return r.stdout


def comparable(report):
    """The report without the lines that time or describe the training itself."""
    keep = []
    for ln in report.splitlines():
        s = ln.strip()
        if s.startswith("ep ") or "steps/s" in s or "trained network" in s:
            continue
        keep.append(ln)
    return keep


def test_report_from_snapshot_matches_full_run(tmp_path):
    snap = str(tmp_path / "a.axs")
    args = KEYS + ["stim=4", "epochs=2", "seed=1"] + sweep.READOUT
    full = cli(args + [f"save={snap}"])
    again = cli(args + [f"load={snap}"])
    a, b = comparable(full), comparable(again)
    assert len(a) > 50
    diff = [(x, y) for x, y in zip(a, b) if x != y]
    assert len(a) == len(b) and not diff, diff[:5]


def test_shared_training_saves_the_cli_network(tmp_path):
    specs = [(4, 0), (6, 1), (4, 2)]
    paths, _ = sweep.train_and_save(KEYS, specs, 2, str(tmp_path), lambda m: None)
    for (n, sd), p in zip(specs, paths):
        own = str(tmp_path / f"cli{n}_{sd}.axs")
        cli(KEYS + [f"stim={n}", "epochs=2", f"seed={sd}", "quiet=1"] + sweep.READOUT + [f"save={own}"])
        x, y = axon_amd.Net.load(p), axon_amd.Net.load(own)
        for b in ("W0", "W1", "W2", "W3", "CONN", "BIAS"):
            assert axon_amd.checksum(x, b) == axon_amd.checksum(y, b), (n, sd, b)
        x.close()
        y.close()


def test_sweep_table(tmp_path, capsys):
    sweep.main(["--table", "seeds", "--stim", "4,6", "--seeds", "0,1", "--epochs", "2", "--keys", " ".join(SMALL),
                "--out", str(tmp_path)])
    out = capsys.readouterr().out
    rows = [ln for ln in out.splitlines() if ln.strip().startswith("NSTIM=")]
    assert len(rows) == 2 and "nan" not in out, out
    for n in (4, 6):
        for sd in (0, 1):
            assert os.path.getsize(tmp_path / f"stim{n}_seed{sd}.txt") > 1000
