"""The workloads at their stated full size (131,072 neurons x 8,192 fan-in per
network), pinned to numbers the reference itself publishes where it has them.

* Table 1 (seed 0, all ten loads 12-300) and Table 3 (seed 0, shared vs exclusive
  members, loads 12-60 and the untrained load-48 control) of
  paper/axon_preprint.tex:259-272 and :333-345: the drop-in CLI is run with the
  reference scripts' keys (bench/cap10.sh, bench/table3_subgroups.sh) and its
  report is parsed the way those scripts parse it.  One run per load carries
  both scripts' readout keys; the readouts only present without learning
  (src/axon.mm:1601-1801), so each table's numbers are those of its own run.
* 25 networks (loads 12-60 x seeds 0-4) trained in shared launches: every
  network's planes, connectivity and bias equal those of the same network
  trained alone, which the golden-row and parity tests tie to the reference.
* 256 recall probes on the trained load-60 network: every probe tile gives
  each probe's spike history of that probe presented alone.
* Serving 8 resident networks at loads 12-200 with free runs up to 200 steps:
  every query's result in the mix equals the query run alone."""
import json
import os
import re

import numpy as np
import pytest

import axon_amd
from conftest import CONFIG_C
from test_gpu_golden import cap10_fields, run_cli

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PUB = json.load(open(os.path.join(HERE, "golden", "published.json")))


def split_fields(out):
    """The three lines bench/table3_subgroups.sh prints from the SHARED vs
    EXCLUSIVE block: shared fraction, shared and exclusive density (%)."""
    frac = float(re.search(r"shared fraction of assembly : ([0-9.]+)", out).group(1))
    ds = float(re.search(r"within-assembly density  SHARED +([0-9.]+)%", out).group(1))
    de = float(re.search(r"within-assembly density  EXCLUSIVE ([0-9.]+)%", out).group(1))
    return frac, ds, de


def close(x, published, decimals):
    """x agrees with a value the paper prints to `decimals` places."""
    return abs(float(x) - published) <= 0.5 * 10.0 ** -decimals + 1e-9


@pytest.mark.parametrize("row", PUB["table1"], ids=lambda r: f"stim{r['stim']}")
def test_table1_and_table3_rows_seed0(row):
    n = row["stim"]
    out = run_cli(CONFIG_C + [f"stim={n}", "epochs=40", "seed=0", "deep=1", "wave=1", "wavet=35", "split=1"])
    w, s, o, m, sus, d = cap10_fields(out)
    got = dict(within=w, recall=s, other=o, margin=m, sustained=sus, drive=d)
    print(f"stim {n}: {got}", flush=True)
    for key, dec in (("within", 2), ("recall", 3), ("other", 3), ("margin", 3), ("sustained", 3)):
        assert close(got[key], row[key], dec), (n, key, got[key], row[key])
    assert int(d) == row["drive"], (n, d, row["drive"])
    t3s = [r for r in PUB["table3"] if r["stim"] == n and r["epochs"] == 40]
    if not t3s:  # Table 3 covers loads 12-60
        return
    t3 = t3s[0]
    frac, ds, de = split_fields(out)
    print(f"stim {n}: shared fraction {frac}, density shared {ds}%, exclusive {de}%", flush=True)
    assert close(frac, t3["shared_fraction"], 3) and close(ds, t3["dens_shared"], 2) and \
        close(de, t3["dens_exclusive"], 2), (n, frac, ds, de, t3)


def test_untrained_control_load48_table1_and_table3():
    out = run_cli(CONFIG_C + ["stim=48", "epochs=0", "seed=0", "deep=1", "wave=1", "wavet=35", "split=1"])
    w, s, o, m, sus, d = cap10_fields(out)
    assert w == PUB["untrained"]["within_pct"] and d == PUB["untrained"]["drive_t6"]
    lo, hi = PUB["untrained"]["margin_range"]
    assert lo - 5e-4 <= float(m) <= hi + 5e-4, m
    t3 = [r for r in PUB["table3"] if r["epochs"] == 0][0]
    frac, ds, de = split_fields(out)
    assert close(frac, t3["shared_fraction"], 3) and close(ds, t3["dens_shared"], 2) and \
        close(de, t3["dens_exclusive"], 2), (frac, ds, de, t3)


def _table(name, tmp_path):
    from axon_amd import sweep

    loads, seeds = sweep.TABLES[name][0], sweep.TABLES[name][1]
    outs = sweep.run_table(name, CONFIG_C, loads, seeds, 40, str(tmp_path), lambda m: print(m, flush=True))
    lines, rows = sweep.format_table(name, loads, seeds, 40, outs)
    print("\n".join(lines), flush=True)
    return rows


def test_table2_five_seeds_networks_trained_together(tmp_path):
    """bench/seeds.sh over loads 24-60 x seeds 0-4 (paper:284-293): 25 networks
    trained in shared launches, each reported in-process."""
    rows = _table("seeds", tmp_path)
    for pub in PUB["table2"]:
        r = rows[pub["stim"]]
        for key, dec in (("sustain_mean", 3), ("sustain_sd", 3), ("margin_mean", 3), ("margin_sd", 3),
                         ("within_mean", 2), ("within_sd", 2)):
            assert close(r[key], pub[key], dec), (pub["stim"], key, r[key], pub[key])
        # the paper's CV is the printed SD over the printed mean (3 decimals each)
        cv = 100.0 * round(r["sustain_sd"], 3) / round(r["sustain_mean"], 3)
        assert close(cv, pub["cv_pct"], 1), (pub["stim"], cv, r)


def test_table4_loop_gain_five_seeds(tmp_path):
    """bench/table4_gaintraj.sh over loads 12, 36, 60 x seeds 0-4 (paper:391-399)."""
    rows = _table("table4", tmp_path)
    for pub in PUB["table4"]:
        for got, want in zip(rows[pub["stim"]], pub["gain"]):
            assert close(got, want, 3), (pub["stim"], rows[pub["stim"]], pub["gain"])


def test_table5_visited_population_five_seeds(tmp_path):
    """bench/table5_union.sh over loads 12, 36, 60 x seeds 0-4 and the untrained
    load-36 control over seeds 0-2 (paper:454-463)."""
    rows = _table("table5", tmp_path)
    for pub in PUB["table5"]:
        m = rows[(pub["stim"], pub["epochs"])]
        got = dict(in_asm=m[0], outside=m[1], dens_in=m[2], dens_out=m[3], dens_all=m[4], mult_out=m[5],
                   mult_pool=m[6], pct_outside=m[7])
        for key, dec in (("in_asm", 0), ("outside", 0), ("pct_outside", 1), ("dens_in", 2), ("dens_out", 2),
                         ("dens_all", 2), ("mult_out", 2), ("mult_pool", 2)):
            assert close(got[key], pub[key], dec), (pub["stim"], pub["epochs"], key, got[key], pub[key])


def train_alone(net):
    """The CLI's training: epochs x patterns, then the extra presentation of
    pattern 0 (src/axon.mm:1009-1010, 1033)."""
    c = net.config
    for _ in range(c.epochs):
        for p in range(c.stim):
            net.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p, outputs=False)
    net.present(0, 1.0, c.T, c.trainclamp, True, False, 31337, outputs=False)
    net.sync()


STATE = ("W0", "W1", "W2", "W3", "CONN", "BIAS")


def test_25_networks_batched_equal_plain():
    loads, seeds = (12, 24, 36, 48, 60), range(5)
    specs = [(P, sd) for P in loads for sd in seeds]
    mk = lambda P, sd: axon_amd.Config.parse(CONFIG_C + [f"stim={P}", "epochs=40", f"seed={sd}"])  # noqa: E731
    ref = {}
    for P, sd in specs:
        net = axon_amd.Net(mk(P, sd))
        train_alone(net)
        ref[(P, sd)] = [axon_amd.checksum(net, b) for b in STATE]
        net.close()
        print(f"plain network stim={P} seed={sd} trained", flush=True)
    nets = [axon_amd.Net(mk(P, sd)) for P, sd in specs]
    mb = axon_amd.MultiNet(nets)
    ms = mb.train([40] * len(nets), extra=True)
    print(f"25 networks trained together in {ms:.0f} ms", flush=True)
    mb.close()
    for net, spec in zip(nets, specs):
        assert [axon_amd.checksum(net, b) for b in STATE] == ref[spec], spec
        net.close()


def probe_set(net):
    """60 stored patterns x {0, 10, 25, 50 %} of their 64 cue columns replaced
    by other input columns, plus 16 random 64-column cues (256 probes)."""
    cues = [axon_amd.probe_cue(net, p, r, 100000 + 1000 * r + p) for r in (0, 6, 16, 32)
            for p in range(net.config.stim)]
    cues += [axon_amd.probe_cue(net, -1, 0, 200000 + i) for i in range(16)]
    return np.stack(cues)


def test_256_probes_on_trained_load60_network_every_tile():
    net = axon_amd.Net(axon_amd.Config.parse(CONFIG_C + ["stim=60", "epochs=40", "seed=0"]))
    train_alone(net)
    stims = probe_set(net)
    assert stims.shape == (256, 4096)
    T, clamp = 35, 5
    alone = [net.present(0, 1.0, T, clamp, False, True, 0, 0, stim=s) for s in stims]
    for tile in (1, 2, 4, 8, 16, 32, 64):
        pb = axon_amd.Probes(net, len(stims), T, tile)
        print(f"256 probes, tile {tile}", flush=True)
        spk, hist, _ = pb.run(stims, T, clamp, 0)
        pb.close()
        for i, r in enumerate(alone):
            assert np.array_equal(hist[i], r.hist), (tile, i)
            assert np.array_equal(spk[i], r.spikes), (tile, i)
    net.close()


def test_serving_8_resident_networks_up_to_200_steps_alone_equals_mix():
    loads = (12, 36, 60, 100, 140, 170, 200, 200)
    nets = [axon_amd.Net(axon_amd.Config.parse(CONFIG_C + [f"stim={P}", "epochs=40", f"seed={i}", "maxt=200"]))
            for i, P in enumerate(loads)]
    mb = axon_amd.MultiNet(nets)
    mb.train([40] * len(nets), extra=True)
    mb.close()
    print("8 resident networks trained", flush=True)
    rs = np.random.RandomState(7)
    qs = []
    for i in range(400):
        k = int(rs.randint(0, len(nets)))
        if i % 16 == 15:
            cue = axon_amd.probe_cue(nets[k], -1, 0, 300000 + i)
        else:
            cue = axon_amd.probe_cue(nets[k], int(rs.randint(0, loads[k])), (0, 6, 16, 32)[i % 4], 300000 + i)
        qs.append((k, cue, int(rs.randint(10, 201))))
    srv = axon_amd.Server(nets, slots_per_net=128, max_t=200, tile=4, steps_per_tick=20)
    ids, res, nxt = {}, {}, 0
    while len(res) < len(qs):
        for _ in range(int(rs.randint(0, 40))):  # bursts join while earlier queries are mid-flight
            if nxt < len(qs):
                k, cue, T = qs[nxt]
                ids[srv.submit(k, cue, T, 5, 0)] = nxt
                nxt += 1
        if srv.tick():
            for rid in srv.completed():
                res[ids.pop(rid)] = srv.result(rid, 200)
    srv.close()
    print("400 queries served", flush=True)
    for i, (k, cue, T) in enumerate(qs):
        h = nets[k].present(0, 1.0, T, 5, False, True, 0, 0, stim=cue)
        assert np.array_equal(res[i][0], h.spikes), i
        assert np.array_equal(res[i][1], axon_amd.row_digest(h.hist)), i
    for n in nets:
        n.close()
