"""Batched sweep: a CLI run that starts from a snapshot (load=) prints the same
report as the full run that saved it (save=), apart from the training lines;
the in-process report (axon_run) on networks trained together in shared
launches is the CLI's full-run report; the sweep driver's tables parse those
reports."""
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


def test_snapshot_reused_for_other_readouts(tmp_path):
    """A snapshot saved by a plain run serves a run with other readout keys
    (split, unionchar, diag): its report equals the full run with those keys;
    a training key that differs is refused."""
    snap = str(tmp_path / "b.axs")
    base = KEYS + ["stim=5", "epochs=2", "seed=4"]
    cli(base + ["quiet=1", f"save={snap}"])
    extra = ["split=1", "unionchar=1", "diag=1", "deep=1", "wave=1", "wavet=30"]
    full = comparable(cli(base + extra))
    again = comparable(cli(base + extra + [f"load={snap}"]))
    diff = [(x, y) for x, y in zip(full, again) if x != y]
    assert len(full) == len(again) and not diff, diff[:5]
    assert any(ln.startswith("UNIONCHAR") for ln in again)
    r = subprocess.run([sweep.AXON_BIN] + base + ["biasup=70", f"load={snap}"], capture_output=True, text=True)
    assert r.returncode == 2 and "training keys" in r.stderr


def test_report_on_networks_trained_together_is_the_cli_report(tmp_path):
    specs = [(4, 0, 2), (6, 1, 2), (4, 2, 2), (5, 3, 0)]
    nets, _ = sweep.train_together(KEYS, specs, lambda m: None)
    extra = ["split=1", "unionchar=1", "diag=1"] + sweep.READOUT
    for spec, net in zip(specs, nets):
        n, sd, ep = spec
        got = comparable(sweep.report(KEYS, spec, extra, net, str(tmp_path)))
        want = comparable(cli(KEYS + [f"stim={n}", f"epochs={ep}", f"seed={sd}"] + extra))
        diff = [(x, y) for x, y in zip(want, got) if x != y]
        assert len(got) == len(want) and not diff, (spec, diff[:5])
        net.close()


@pytest.mark.parametrize("table", sorted(sweep.TABLES))
def test_sweep_tables_small(table, tmp_path, capsys):
    loads = {"cap10": "4,6", "seeds": "4,6", "table3": "4,6", "table4": "4,6", "table5": "4,6"}[table]
    sweep.main(["--table", table, "--stim", loads, "--seeds", "0,1", "--epochs", "2", "--keys", " ".join(SMALL),
                "--out", str(tmp_path)])
    out = capsys.readouterr().out
    assert "nan" not in out, out
    print(out)
    for n in (4, 6):
        assert os.path.getsize(tmp_path / f"stim{n}_seed0_ep2.txt") > 1000
