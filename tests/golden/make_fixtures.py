"""Writes the oracle-derived golden fixtures used by the GPU tests.

Runs the CPU oracle (oracle/axon_oracle) on the full-scale golden-row
configuration, on its untrained control and on three small (ncol=512)
configurations, each with every report block switched on (deep, wave, diag,
split, unionchar), and stores the full report (every line the oracle
restates; the GPU tests look for each of them, in order, in the CLI's
output).  Usage: python tests/golden/make_fixtures.py
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
import oracle  # noqa: E402
from conftest import CONFIG_C  # noqa: E402

BLOCKS = "diag=1 split=1 unionchar=1".split()
RUNS = {
    "c1_trained": CONFIG_C + "stim=12 epochs=40 deep=1 wave=1 wavet=35 seed=0".split() + BLOCKS,
    "c1_untrained": CONFIG_C + "stim=12 epochs=0 deep=1 wave=1 wavet=35 seed=0".split() + BLOCKS,
    "c0_small": CONFIG_C + "ncol=512 incols=128 stim=12 epochs=10 deep=1 wave=1 wavet=35 seed=0".split() + BLOCKS,
    # the reference's built-in defaults (legacy rule, margin gate, trainclamp = T), small
    "c0_defaults": "ncol=512 incols=128 stimcols=64 stim=8 epochs=4 deep=1 wave=1 wavet=30 seed=3".split() + BLOCKS,
    # Configuration C with longer presentations (deep trace 40 steps, 20-step tables)
    "c0_t20": CONFIG_C + "ncol=512 incols=128 stim=6 epochs=3 T=20 deep=1 wave=1 wavet=45 seed=1".split() + BLOCKS,
}


def main():
    oracle.build()
    out = {}
    for name, args in RUNS.items():
        txt = subprocess.run([oracle.CLI_PATH] + args + ["threads=8"], capture_output=True, text=True,
                             check=True).stdout
        lines = txt.splitlines()
        pinned = [l for l in lines if l.startswith("PINNED")][0]
        out[name] = {"args": args, "pinned": dict(kv.split("=") for kv in pinned.split()[1:]),
                     "report": [l for l in lines if not l.startswith("PINNED")]}
        print(name, pinned, flush=True)
    with open(os.path.join(HERE, "oracle_reports.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
