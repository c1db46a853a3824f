"""GPU unit tests of the HIP kernel building blocks against the oracle and closed
forms: bit-serial counters (k_bstest, mm:424-434), k-of-32 winners-take-all
(mm:134-144) and the column top-K (mm:259-352).  Integer work: bit-exact."""
import numpy as np
import pytest

import axon_amd
import oracle

pytestmark = pytest.mark.gpu


def test_bitserial_ops_match_closed_forms():
    # the reference's vbs generator: mt19937(424242), v = r()&15 then m = r()&1 per synapse
    rng = np.random.default_rng(424242)
    nsyn = (1 << 14) * 128
    v = rng.integers(0, 16, nsyn)
    m = rng.integers(0, 2, nsyn)
    planes = [np.packbits(((v >> b) & 1).astype(np.uint8), bitorder="little").view(np.uint32) for b in range(4)]
    mask = np.packbits(m.astype(np.uint8), bitorder="little").view(np.uint32)

    def unpack(p):
        bits = [np.unpackbits(x.view(np.uint8), bitorder="little") for x in p]
        return bits[0] + 2 * bits[1] + 4 * bits[2] + 8 * bits[3]

    for op, lo, hi, ref in [(0, 0, 0, np.where(m == 1, np.minimum(15, v + 1), v)),
                            (1, 0, 0, np.where(m == 1, np.maximum(0, v - 1), v)),
                            (2, 5, 11, np.clip(v, 5, 11)), (2, 6, 10, np.clip(v, 6, 10)),
                            (2, 4, 12, np.clip(v, 4, 12))]:
        print(f"bs op={op} lo={lo} hi={hi}", flush=True)
        out = axon_amd.bs_test(op, lo, hi, planes, mask)
        assert np.array_equal(unpack(out), ref), (op, lo, hi)
        # and word-for-word against the oracle restatement
        ora = oracle.bs_apply(op, lo, hi, planes, mask)
        for a, b in zip(out, ora):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("k", list(range(0, 35)))
def test_kwta_matches_oracle(k):
    rng = np.random.default_rng(100 + k)
    rows = []
    for hi in (2, 3, 8, 200, 1 << 20, 1 << 31):
        rows.append(rng.integers(0, hi, (64, 32)))
    rows.append(np.zeros((4, 32), np.int64))
    rows.append(np.full((4, 32), 0xFFFFFFFF, np.int64))
    v = np.concatenate(rows).astype(np.uint32)
    print(f"kwta k={k} rows={v.shape[0]}", flush=True)
    ball, vk, vk1 = axon_amd.kwta_test(v, k)
    for i in range(v.shape[0]):
        assert (int(ball[i]), int(vk[i]), int(vk1[i])) == oracle.kwta(v[i], k), (k, i, v[i].tolist())


@pytest.mark.parametrize("par", [True, False])
def test_topk_matches_oracle(par):
    rng = np.random.default_rng(7 if par else 8)
    cases = []
    for trial in range(96):
        ncol, incols = [(512, 128), (4096, 512), (1024, 8)][trial % 3]
        scale = [2, 40, 4000, 1 << 20][trial % 4]
        d = rng.integers(0, scale, ncol).astype(np.uint32)
        target = int(rng.integers(0, ncol - incols + 40))  # includes 0 and > candidates (degenerate)
        cases.append((d, ncol, incols, target))
    for d, ncol, incols, target in cases:
        th = axon_amd.topk_test(d, incols, target, par)[0]
        assert (int(th[0]), int(th[1])) == oracle.topk(d, incols, target, par), (ncol, incols, target)


def test_topk_batched_instances():
    rng = np.random.default_rng(9)
    d = rng.integers(0, 3000, (37, 4096)).astype(np.uint32)
    th = axon_amd.topk_test(d, 512, 64, True)
    for i in range(d.shape[0]):
        assert tuple(int(x) for x in th[i]) == oracle.topk(d[i], 512, 64, True)


def test_mem_probe_runs_and_reports_plausible_bandwidth():
    """The calibration probes run (each shape, a small and a cache-exceeding
    size) and report a positive rate below the 8 TB/s peak."""
    for kind in (0, 1, 2):
        for nbytes in (64 << 20, 1 << 30):
            g = axon_amd.mem_probe(kind, nbytes, 2)
            print(f"mem probe kind {kind} {nbytes >> 20} MiB: {g:.0f} GB/s", flush=True)
            assert 100.0 < g < 8000.0


def test_sync_probe_orders_the_synchronisation_costs():
    """The synchronisation probe runs every kind at a propagation launch's
    shape; a per-workgroup agent-scope release fence costs more than the
    stores alone (the basis for separate launches per timestep phase), and
    bad arguments are refused."""
    us = {kind: axon_amd.sync_probe(kind, 2, 256, 20) for kind in range(4)}
    print(f"sync probe 2x256: {us}", flush=True)
    assert all(0.0 < v < 10000.0 for v in us.values())
    assert us[2] > us[0]
    with pytest.raises(axon_amd.AxonError):
        axon_amd.sync_probe(4, 1, 1, 1)
    with pytest.raises(axon_amd.AxonError):
        axon_amd.sync_probe(0, 0, 1, 1)


def test_prop_probe_reads_exactly_the_spiking_columns_lines():
    """The propagation probe runs the network's propagation on the given
    activity: a launch reads one W3 line per (tile, spiking presynaptic
    column of its group), dense activity reads every line, none reads none,
    and the planes are left as they were."""
    net = axon_amd.Net(axon_amd.Config.parse(["ncol=512", "incols=128", "stim=2", "seed=4"]))
    conn = net.read("CONN").astype(np.int64)
    w3 = net.read("W3").copy()
    tiles_per_group = np.bincount(conn, minlength=512 // 8)
    rng = np.random.default_rng(9)

    def host_lines(spk):
        gm = (spk != 0).reshape(-1, 8).sum(axis=1)  # spiking columns per presynaptic group
        return int((gm * tiles_per_group).sum())

    us, lines = axon_amd.prop_probe(net, np.full(512, 0xFFFFFFFF, np.uint32), 3)
    assert lines == 512 * 32 * 8 and us > 0
    assert axon_amd.prop_probe(net, np.zeros(512, np.uint32), 3)[1] == 0
    for frac in (0.01, 0.1, 0.5):
        spk = np.zeros(512, np.uint32)
        cols = rng.choice(512, int(512 * frac), replace=False)
        spk[cols] = rng.integers(1, 1 << 32, len(cols), dtype=np.uint64).astype(np.uint32)
        assert axon_amd.prop_probe(net, spk, 2)[1] == host_lines(spk), frac
    assert np.array_equal(net.read("W3"), w3)
    with pytest.raises(ValueError):
        axon_amd.prop_probe(net, np.zeros(100, np.uint32))
    net.close()
