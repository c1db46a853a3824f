"""Writes the oracle-derived golden fixtures used by the GPU tests.

Runs the CPU oracle (oracle/axon_oracle) on the full-scale golden-row
configuration and on its untrained control, and stores the report lines the
reference's bench scripts parse.  Usage: python tests/golden/make_fixtures.py
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

RUNS = {
    "c1_trained": CONFIG_C + "stim=12 epochs=40 deep=1 wave=1 wavet=35 seed=0".split(),
    "c1_untrained": CONFIG_C + "stim=12 epochs=0 deep=1 wave=1 wavet=35 seed=0".split(),
    "c0_small": CONFIG_C + "ncol=512 incols=128 stim=12 epochs=10 deep=1 wave=1 wavet=35 seed=0".split(),
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
