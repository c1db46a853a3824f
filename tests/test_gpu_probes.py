"""Probe batches (many recall presentations against one network in shared
launches) must give, for every probe, exactly the spike history of that probe
presented alone: the single-network engine and the CPU oracle are the checks.
Covers every probe tile size, probes of different lengths sharing launches,
slots refilled while other probes are mid-flight, bias on/off and the readout
knobs ktest/facil."""
import numpy as np
import pytest

import axon_amd
import oracle
from conftest import C0, CONFIG_C

pytestmark = pytest.mark.gpu

SMALL = ["ncol=512", "incols=128"]


def trained(args, epochs):
    net = axon_amd.Net(axon_amd.Config.parse(args))
    ora = oracle.OracleNet(oracle.parse(args), threads=4)
    cfg = net.config
    for ep in range(epochs):
        for p in range(cfg.stim):
            net.present(p, 1.0, cfg.T, cfg.trainclamp, True, False, 1000 + p, outputs=False)
            ora.present(p, 1.0, cfg.T, cfg.trainclamp, True, 1000 + p)
    net.sync()
    return net, ora


def cues(net, n):
    cfg = net.config
    out = []
    for i in range(n):
        if i % 5 == 4:
            out.append(axon_amd.probe_cue(net, -1, 0, 900 + i))
        else:
            r = [0, 6, 16, 32][i % 4]
            out.append(axon_amd.probe_cue(net, i % cfg.stim, r, 900 + i))
    return np.stack(out)


def test_cue_generator():
    net = axon_amd.Net(axon_amd.Config.parse(C0))
    for p in range(12):
        members = set(net.stim_cols(p).tolist())
        for r in (0, 6, 16, 32, 64):
            c = axon_amd.probe_cue(net, p, r, 77 + r)
            on = set(np.flatnonzero(c).tolist())
            assert all(c[x] == 0xFFFFFFFF for x in on)
            assert len(on) == 64 and len(on & members) == 64 - r and max(on) < 128
        a = axon_amd.probe_cue(net, p, 16, 5)
        assert np.array_equal(a, axon_amd.probe_cue(net, p, 16, 5))  # deterministic in the seed
    rnd = axon_amd.probe_cue(net, -1, 0, 3)
    assert np.count_nonzero(rnd) == 64 and np.flatnonzero(rnd).max() < 128


@pytest.mark.parametrize("tile", [1, 2, 4, 8, 16, 32, 64])
def test_probe_batch_matches_single_presentations(tile):
    net, ora = trained(C0, 3)
    stims = cues(net, 40)
    T, clamp = 35, 5
    pb = axon_amd.Probes(net, slots=40, max_t=64, tile=tile)
    print(f"probe tile {tile}: {pb.slots} slots", flush=True)
    spk, hist, ms = pb.run(stims, T, clamp, use_bias=0)
    for i in range(stims.shape[0]):
        r = net.present(0, 1.0, T, clamp, False, True, 0, 0, stim=stims[i])
        assert np.array_equal(hist[i], r.hist), f"tile {tile} probe {i}: history"
        assert np.array_equal(spk[i], r.spikes), f"tile {tile} probe {i}: final"
        if i % 7 == 0:
            ora.present(0, 1.0, T, clamp, False, 0, 0, stim=stims[i])
            assert np.array_equal(hist[i].ravel(), ora.buf("HIST")[: T * 512]), f"probe {i}: oracle"


def test_probes_of_different_lengths_and_refilled_slots():
    net, ora = trained(C0, 2)
    stims = cues(net, 48)
    rng = np.random.default_rng(11)
    Ts = rng.integers(1, 64, 48).astype(np.uint32)
    pb = axon_amd.Probes(net, slots=16, max_t=64, tile=8)
    results = {}
    pending = list(range(48))
    inflight = {}
    # first fill
    for s in range(16):
        i = pending.pop(0)
        pb.load(s, stims[i][None], [Ts[i]], 5, use_bias=1)
        inflight[s] = i
    while inflight:
        pb.step(7)
        spk, hist, done = pb.read(0, 16)
        for s, i in list(inflight.items()):
            if done[s] >= Ts[i]:
                results[i] = (spk[s].copy(), hist[s, : Ts[i]].copy())
                del inflight[s]
                if pending:  # refill this slot while the others are mid-flight
                    j = pending.pop(0)
                    pb.load(s, stims[j][None], [Ts[j]], 5, use_bias=1)
                    inflight[s] = j
    assert len(results) == 48
    for i, (spk, hist) in results.items():
        r = net.present(0, 1.0, int(Ts[i]), 5, False, True, 0, 1, stim=stims[i])
        assert np.array_equal(hist, r.hist), f"probe {i} T={Ts[i]}"
        assert np.array_equal(spk, r.spikes)


def test_probe_readout_knobs():
    args = SMALL + CONFIG_C + ["ktest=10", "facil=40", "stim=6", "seed=2"]
    net, ora = trained(args, 2)
    stims = np.stack([axon_amd.probe_cue(net, p % 6, 8 * (p % 3), p) for p in range(24)])
    pb = axon_amd.Probes(net, slots=24, max_t=40, tile=16)
    spk, hist, _ = pb.run(stims, 30, 5, use_bias=0)
    for i in range(24):
        ora.present(0, 1.0, 30, 5, False, 0, 0, stim=stims[i])
        assert np.array_equal(hist[i].ravel(), ora.buf("HIST")[: 30 * 512]), f"probe {i}"


def test_probe_errors():
    net = axon_amd.Net(axon_amd.Config.parse(C0))
    with pytest.raises(axon_amd.AxonError):
        axon_amd.Probes(net, slots=8, max_t=16, tile=12)
    pb = axon_amd.Probes(net, slots=8, max_t=16, tile=8)
    with pytest.raises(axon_amd.AxonError) as e:
        pb.load(0, np.zeros((1, 512), np.uint32), [17], 5)
    assert e.value.status == 6
    with pytest.raises(axon_amd.AxonError):
        pb.load(7, np.zeros((2, 512), np.uint32), [5, 5], 5)


@pytest.mark.parametrize("tile", [1, 4, 16])
def test_occupancy_limits_change_no_result(tile):
    """Capping the propagation's waves per CU (LDS padding) changes only timing."""
    net, _ = trained(C0, 2)
    stims = cues(net, 24)
    T, clamp = 30, 5
    ref = None
    for waves in axon_amd.PROBE_WAVES:
        pb = axon_amd.Probes(net, slots=24, max_t=64, tile=tile, waves_per_cu=waves)
        spk, hist, _ = pb.run(stims, T, clamp, use_bias=0)
        pb.close()
        if ref is None:
            ref = (spk, hist)
        else:
            assert np.array_equal(spk, ref[0]) and np.array_equal(hist, ref[1]), f"tile {tile} waves {waves}"


def test_autotuner_sweeps_tiles_and_occupancies():
    net, _ = trained(C0, 2)
    stims = cues(net, 16)
    best, waves, table = axon_amd.probe_autotune(net, stims, 20, 5, 0)
    assert best in axon_amd.PROBE_TILES and waves in axon_amd.PROBE_WAVES
    assert len(table) == len(axon_amd.PROBE_TILES) * len(axon_amd.PROBE_WAVES)
    assert all(v > 0 for v in table.values()) and table[(best, waves)] == min(table.values())
    with pytest.raises(axon_amd.AxonError):
        axon_amd.Probes(net, slots=8, max_t=32, tile=4, waves_per_cu=24)
    # the run went into the tuning table under (ncol, probes, T)
    e = axon_amd.tune_lookup(net.ncol, 16, 20)
    assert (e["probes"], e["T"], e["tile"], e["waves_per_cu"], e["reps"]) == (16, 20, best, waves, 3)
    assert e["ms"] == table[(best, waves)] and e["spread_ms"] >= 0


def test_autotuner_batch_sizes_and_server_choice(tmp_path):
    """The batch-size axis gives one table entry per size; a server created
    with tile 0 takes the nearest entry, and its results stay those of
    single presentations; a saved table restores the same choice."""
    net, _ = trained(C0, 2)
    stims = cues(net, 32)
    axon_amd.tune_clear()
    got = axon_amd.probe_autotune_sizes(net, stims, 24, 5, [4, 16, 32])
    assert [e["probes"] for e in got] == [4, 16, 32]
    assert all(e["ncol"] == net.ncol and e["T"] == 24 and e["reps"] == 3 and e["ms"] > 0 for e in got)
    assert all(e["tile"] in axon_amd.PROBE_TILES and e["waves_per_cu"] in axon_amd.PROBE_WAVES for e in got)
    assert len(axon_amd.tune_entries()) == 3
    with pytest.raises(axon_amd.AxonError):
        axon_amd.probe_autotune_sizes(net, stims, 24, 5, [33])
    f = tmp_path / "tune.txt"
    axon_amd.tune_save(str(f))
    axon_amd.tune_clear()
    axon_amd.tune_load(str(f))
    srv = axon_amd.Server([net], slots_per_net=16, max_t=24, tile=0, steps_per_tick=6)
    cfg = srv.config()
    want = [e for e in got if e["probes"] == 16][0]
    assert (cfg["tile"], cfg["waves_per_cu"]) == (want["tile"], want["waves_per_cu"])
    ids = {srv.submit(0, stims[i], 24): i for i in range(20)}
    results = {}
    while len(results) < 20:
        srv.tick()
        for rid in srv.completed():
            results[ids[rid]] = srv.result(rid, 24)
    for i, (spk, dig, _, T) in results.items():
        r = net.present(0, 1.0, 24, 5, False, True, 0, 0, stim=stims[i])
        assert np.array_equal(spk, r.spikes) and np.array_equal(dig, axon_amd.row_digest(r.hist)), f"query {i}"
    srv.close()
    axon_amd.tune_clear()
    srv = axon_amd.Server([net], slots_per_net=16, max_t=24, tile=0)  # empty table: tile 4, 32 waves per CU
    assert (srv.config()["tile"], srv.config()["waves_per_cu"]) == (4, 32)
    srv.close()


@pytest.mark.parametrize("mode_args", [["mode=0"], ["mode=1", "dthr=2", "boost=3"], ["mode=2", "dthr=2"],
                                       ["mode=1", "boost=3000"]],
                         ids=["mode0", "mode1-dthr2", "mode2", "mode1-boost3000-acc32"])
def test_tile_walks_and_dendritic_modes(mode_args):
    """130 probes at ncol=512: tiles 1 and 2 walk several tiles per wave (the last run ragged), tiles 4 and 8
    one; every mode of the branch output; all equal to the per-branch kernel (tile 16), to single
    presentations and, for a sample, to the oracle.  A boost of 3000 lets accumulators pass 2^16, so the
    probes keep 32-bit accumulators; the other cases store 16-bit ones."""
    args = C0 + mode_args
    net, ora = trained(args, 1)
    stims = cues(net, 130)
    T, clamp = 24, 5
    res = {}
    for tile in (16, 1, 2, 4, 8):
        pb = axon_amd.Probes(net, slots=130, max_t=32, tile=tile)
        print(f"{mode_args} tile {tile}", flush=True)
        res[tile] = pb.run(stims, T, clamp, use_bias=1)[:2]
        pb.close()
    for tile in (1, 2, 4, 8):
        assert np.array_equal(res[tile][0], res[16][0]), f"{mode_args} tile {tile}: final spikes"
        assert np.array_equal(res[tile][1], res[16][1]), f"{mode_args} tile {tile}: histories"
    for i in range(0, 130, 13):
        r = net.present(0, 1.0, T, clamp, False, True, 0, 1, stim=stims[i])
        assert np.array_equal(res[1][1][i], r.hist), f"{mode_args} probe {i}: single presentation"
        if i % 39 == 0:
            ora.present(0, 1.0, T, clamp, False, 0, 1, stim=stims[i])
            assert np.array_equal(res[1][1][i].ravel(), ora.buf("HIST")[: T * 512]), f"{mode_args} probe {i}: oracle"


def test_wide_network_takes_the_per_branch_kernel():
    """Item words hold spike indices below 2^20: at ncol = 131080 a tile of 8 probes runs the per-branch
    kernel instead, with the same results as tile 1 (items) and as single presentations."""
    args = ["ncol=131080", "incols=512", "stim=2", "seed=5"] + CONFIG_C
    net = axon_amd.Net(axon_amd.Config.parse(args))
    stims = np.stack([axon_amd.probe_cue(net, i % 2, 8 * i, 40 + i) for i in range(8)])
    T, clamp = 8, 4
    out = {}
    for tile in (1, 8):
        pb = axon_amd.Probes(net, slots=8, max_t=16, tile=tile)
        print(f"ncol 131080 tile {tile}", flush=True)
        out[tile] = pb.run(stims, T, clamp, use_bias=1)[:2]
        pb.close()
    assert np.array_equal(out[1][0], out[8][0]) and np.array_equal(out[1][1], out[8][1])
    for i in (0, 5):
        r = net.present(0, 1.0, T, clamp, False, True, 0, 1, stim=stims[i])
        assert np.array_equal(out[8][1][i], r.hist), f"probe {i}"
    assert np.count_nonzero(out[8][1]) > 0
    net.close()


def test_trained_wide_network_at_the_item_index_limit():
    """ncol = 131072 after training: tile 4 (4 x 131072 = 2^19 spike words) takes the item kernel, tile 8
    (exactly 2^20, the item index limit) and tile 16 the per-branch kernel; all equal single
    presentations."""
    args = ["ncol=131072", "incols=512", "stim=2", "seed=3"] + CONFIG_C
    net = axon_amd.Net(axon_amd.Config.parse(args))
    c = net.config
    for p in range(2):
        net.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p, outputs=False)
    net.present(0, 1.0, c.T, c.trainclamp, True, False, 31337, outputs=False)
    net.sync()
    stims = np.stack([axon_amd.probe_cue(net, i % 2, (0, 6, 16, 32)[i % 4], 70 + i) for i in range(12)])
    T, clamp = 10, 5
    out = {}
    for tile in (4, 8, 16):
        pb = axon_amd.Probes(net, slots=12, max_t=16, tile=tile)
        print(f"ncol 131072 trained tile {tile}", flush=True)
        out[tile] = pb.run(stims, T, clamp, use_bias=0)[:2]
        pb.close()
    for tile in (8, 16):
        assert np.array_equal(out[4][0], out[tile][0]) and np.array_equal(out[4][1], out[tile][1]), tile
    for i in (0, 7, 11):
        r = net.present(0, 1.0, T, clamp, False, True, 0, 0, stim=stims[i])
        assert np.array_equal(out[4][1][i], r.hist), f"probe {i}"
    assert np.count_nonzero(out[4][1]) > 0
    net.close()


def test_dense_cues_match_across_tiles():
    """Dense activity (256 clamped input columns, half the recurrent columns winning): long item lists in
    every tile; tiles 1, 4 and 8 equal the per-branch kernel (tile 16) and single presentations."""
    args = ["ncol=512", "incols=320", "stim=3", "seed=9"] + CONFIG_C + ["stimcols=256", "target=96"]
    net = axon_amd.Net(axon_amd.Config.parse(args))
    stims = np.stack([axon_amd.probe_cue(net, i % 3, 16 * (i % 4), 60 + i) for i in range(24)])
    T, clamp = 12, 6
    out = {}
    for tile in (16, 1, 4, 8):
        pb = axon_amd.Probes(net, slots=24, max_t=16, tile=tile)
        print(f"dense cues tile {tile}", flush=True)
        out[tile] = pb.run(stims, T, clamp, use_bias=1)[:2]
        pb.close()
    for tile in (1, 4, 8):
        assert np.array_equal(out[tile][0], out[16][0]) and np.array_equal(out[tile][1], out[16][1]), tile
    for i in (0, 7, 13):
        r = net.present(0, 1.0, T, clamp, False, True, 0, 1, stim=stims[i])
        assert np.array_equal(out[8][1][i], r.hist), f"probe {i}"
    # items of the densest (column, 8-probe tile, step): (probe, branch, spiking presynaptic column) triples
    conn = net.read("CONN").reshape(512, 32)
    hist = out[8][1]
    most = 0
    for t in range(1, T):
        act = (hist[:, t - 1, :] != 0).reshape(24, 64, 8)
        items = act.sum(axis=2)
        for j in range(3):
            most = max(most, int(items[8 * j:8 * j + 8].sum(axis=0)[conn].sum(axis=1).max()))
    print(f"densest 8-probe tile: {most} items", flush=True)
    assert most > 256
