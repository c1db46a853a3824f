"""Serving: queries of uneven length arriving while others run must each give
exactly the result of that query run alone (final ballots and every step's
ballot row, compared through the row checksum), admission must respect the
HBM budget, and a query that can never fit is rejected up front."""
import numpy as np
import pytest

import axon_amd
from conftest import CONFIG_C

pytestmark = pytest.mark.gpu

SMALL = ["ncol=512", "incols=128", "maxt=200"]


def trained_nets():
    nets = []
    for stim, seed in [(4, 0), (8, 1), (12, 2)]:
        n = axon_amd.Net(axon_amd.Config.parse(SMALL + CONFIG_C + [f"stim={stim}", f"seed={seed}"]))
        c = n.config
        for ep in range(2):
            for p in range(stim):
                n.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p, outputs=False)
        n.sync()
        nets.append(n)
    return nets


def make_requests(nets, count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        k = int(rng.integers(0, len(nets)))
        net = nets[k]
        if i % 5 == 4:
            cue = axon_amd.probe_cue(net, -1, 0, 5000 + i)
        else:
            cue = axon_amd.probe_cue(net, int(rng.integers(0, net.config.stim)), [0, 6, 16, 32][i % 4], 5000 + i)
        reqs.append((k, cue, int(rng.integers(10, 201))))
    return reqs


def isolated(net, cue, T):
    r = net.present(0, 1.0, T, 5, False, True, 0, 0, stim=cue)
    return r.spikes, axon_amd.row_digest(r.hist)


def test_results_same_alone_and_in_the_mix():
    nets = trained_nets()
    srv = axon_amd.Server(nets, slots_per_net=16, max_t=200, tile=8, steps_per_tick=7)
    reqs = make_requests(nets, 90, 1)
    ids, pending = {}, list(range(len(reqs)))
    max_inflight, joined_mid_flight = 0, 0
    results = {}
    rng = np.random.default_rng(2)
    while pending or len(results) < len(reqs):
        for _ in range(int(rng.integers(0, 9))):  # a burst of arrivals this tick
            if pending:
                i = pending.pop(0)
                k, cue, T = reqs[i]
                ids[srv.submit(k, cue, T)] = i
        q, f, _ = srv.status()
        srv.tick()
        q2, f2, _ = srv.status()
        max_inflight = max(max_inflight, f)
        if f and q > q2:
            joined_mid_flight += 1
        for rid in list(ids):
            r = srv.result(rid, 200)
            if r is not None:
                results[ids.pop(rid)] = r
    assert max_inflight > 16 and joined_mid_flight > 5
    for i, (k, cue, T) in enumerate(reqs):
        spk, dig, lat, TT = results[i]
        assert TT == T and lat > 0
        s_alone, d_alone = isolated(nets[k], cue, T)
        assert np.array_equal(spk, s_alone), f"request {i}: final ballots"
        assert np.array_equal(dig, d_alone), f"request {i}: step rows"
    srv.close()


def test_budget_admission_and_rejection():
    nets = trained_nets()[:1]
    per_query = lambda T: T * 512 * 4 + 512 * 32 * 4 + 4 * 512 * 4  # noqa: E731
    budget = 3 * per_query(100) + 10
    srv = axon_amd.Server(nets, slots_per_net=16, max_t=200, tile=8, steps_per_tick=5, hbm_budget=budget)
    cue = axon_amd.probe_cue(nets[0], 0, 0, 1)
    with pytest.raises(axon_amd.AxonError) as e:
        axon_amd.Server(nets, slots_per_net=8, max_t=200, tile=8, steps_per_tick=5,
                        hbm_budget=per_query(50)).submit(0, cue, 150)
    assert e.value.status == 3  # AXON_ERR_BUDGET_EXCEEDED
    ids = [srv.submit(0, cue, 100) for _ in range(8)]
    srv.tick()
    q, f, reserved = srv.status()
    assert f == 3 and q == 5 and reserved <= budget
    done = {}
    while len(done) < 8:
        srv.tick()
        _, f, reserved = srv.status()
        assert f <= 3 and reserved <= budget
        for rid in ids:
            if rid not in done:
                r = srv.result(rid, 200)
                if r is not None:
                    done[rid] = r
    ref = isolated(nets[0], cue, 100)
    for r in done.values():
        assert np.array_equal(r[0], ref[0]) and np.array_equal(r[1], ref[1])


def test_launches_cover_up_to_the_highest_busy_slot():
    """Ticks advance only slots below the highest busy one: a long query in a
    high slot keeps running while the low slots empty and refill."""
    nets = trained_nets()[:1]
    net = nets[0]
    srv = axon_amd.Server(nets, slots_per_net=32, max_t=200, tile=8, steps_per_tick=5)
    reqs = [(axon_amd.probe_cue(net, i % 4, 6, 100 + i), 10) for i in range(20)]  # slots 0..19
    reqs.append((axon_amd.probe_cue(net, 1, 16, 999), 200))  # slot 20: long
    ids = {srv.submit(0, cue, T): i for i, (cue, T) in enumerate(reqs)}
    results, ticks, refilled = {}, 0, False
    while len(results) < len(reqs):
        srv.tick()
        ticks += 1
        for rid in list(ids):
            r = srv.result(rid, 200)
            if r is not None:
                results[ids.pop(rid)] = r
        if not refilled and len(results) >= 20:  # only the long query is in flight now
            assert 20 not in results
            for j in range(5):  # these take the low slots again
                cue = axon_amd.probe_cue(net, j % 4, 0, 500 + j)
                reqs.append((cue, 35))
                ids[srv.submit(0, cue, 35)] = len(reqs) - 1
            refilled = True
    assert refilled and ticks >= 40
    for i, (cue, T) in enumerate(reqs):
        spk, dig, _, TT = results[i]
        s_alone, d_alone = isolated(net, cue, T)
        assert TT == T and np.array_equal(spk, s_alone) and np.array_equal(dig, d_alone), f"request {i}"
    srv.close()
