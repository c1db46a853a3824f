"""Pins of the CPU oracle against what fixes each part independently of it.

* mt19937: the C++ standard's known answer.
* hash32 / k_init_w / k_init_conn: numpy evaluation of the lowbias32 closed
  form, plus the counter-value invariant {5,6,9,14} and the 2^-nand density.
* bit-serial inc/dec/clip: the closed forms the reference's vbs check uses
  (min(15,v+1), max(0,v-1), clamp; src/axon.mm:945-947).
* column top-K: the target-th largest value from a sort; serial == parallel
  whenever a crossing bin exists (mm:958-994).
* k-of-32: brute-force rank on small inputs, popcount(ballot) == k.
* whole pipeline: the published golden row and the untrained control
  (README.md:57-66, paper Table 1 and mm:253-255).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import C0, C1, CONFIG_C

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "published.json")))


def lowbias32(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x.astype(np.uint32)


def test_mt19937_known_answer(oracle_mod):
    k = GOLD["mt19937_kat"]
    assert oracle_mod.mt_nth(k["seed"], k["n"]) == k["value"]


def test_uniform01_is_two_word_canonical(oracle_mod):
    # generate_canonical<double,53> over a 32-bit engine: (g1 + g2 * 2^32) / 2^64
    for seed in (0, 1, 4400, 123456789):
        g1, g2 = oracle_mod.mt_nth(seed, 1), oracle_mod.mt_nth(seed, 2)
        expect = (float(g1) + float(g2) * 4294967296.0) / 18446744073709551616.0
        assert oracle_mod.uniform01(seed, 0) == expect
        assert 0.0 <= oracle_mod.uniform01(seed, 3) < 1.0


def test_shuffle_is_a_permutation(oracle_mod):
    for n in (2, 3, 64, 128, 512):
        a = np.arange(n, dtype=np.uint32)
        s = oracle_mod.libcxx_shuffle(a, 7)
        assert sorted(s.tolist()) == list(range(n))
        assert n <= 3 or (s != a).any()


@pytest.fixture(scope="module")
def small_net(oracle_mod):
    return oracle_mod.OracleNet(oracle_mod.parse(C0), threads=1)


def test_init_planes_closed_form(small_net):
    W = [small_net.buf(f"W{i}").copy() for i in range(4)]
    seed, nand = 0, 4
    idx = np.arange(0, W[0].size // 4, 97, dtype=np.uint64)  # uint4 indices
    g = (idx.astype(np.uint32) ^ np.uint32((seed * 0x9E3779B9) & 0xFFFFFFFF)).astype(np.uint64)
    for q in range(4):
        m = np.full(idx.size, 0xFFFFFFFF, np.uint32)
        for i in range(nand):
            m &= lowbias32(g * 32 + i * 4 + q)
        c = lowbias32(g * 8 + (11, 13, 17, 19)[q])
        w = (idx * 4 + q).astype(np.int64)
        assert np.array_equal(W[3][w], m)
        assert np.array_equal(W[1][w], c)
        assert np.array_equal(W[0][w], ~c)
        assert np.array_equal(W[2][w], (~m) | c)


def test_init_counter_values_and_density(small_net):
    W = [small_net.buf(f"W{i}")[: 1 << 16] for i in range(4)]
    bits = np.unpackbits(np.stack([w.view(np.uint8) for w in W]), axis=1, bitorder="little")
    val = bits[0] + 2 * bits[1] + 4 * bits[2] + 8 * bits[3]
    assert set(np.unique(val).tolist()) <= {5, 6, 9, 14}
    assert abs(bits[3].mean() - 2.0 ** -4) < 0.002


def test_init_conn_closed_form(small_net):
    conn = small_net.buf("CONN").copy()
    seed, inbr, ngrp, ningrp = 0, 4, 512 // 8, 128 // 8
    g = np.arange(conn.size, dtype=np.uint64)
    h = lowbias32((g * 2246822519 + 7 + seed * 2654435761) & 0xFFFFFFFF).astype(np.int64)
    s = g % 32
    expect = np.where(s < inbr, h % ningrp, ningrp + h % (ngrp - ningrp))
    assert np.array_equal(conn, expect.astype(np.uint32))


def test_bitserial_closed_forms(oracle_mod):
    rng = np.random.default_rng(424242)
    n = 4096
    v = rng.integers(0, 16, n * 32)
    m = rng.integers(0, 2, n * 32)
    planes = [np.packbits(((v >> b) & 1).astype(np.uint8), bitorder="little").view(np.uint32) for b in range(4)]
    mask = np.packbits(m.astype(np.uint8), bitorder="little").view(np.uint32)

    def unpack(p):
        bits = [np.unpackbits(x.view(np.uint8), bitorder="little") for x in p]
        return bits[0] + 2 * bits[1] + 4 * bits[2] + 8 * bits[3]

    for op, lo, hi, ref in [(0, 0, 0, np.where(m == 1, np.minimum(15, v + 1), v)),
                            (1, 0, 0, np.where(m == 1, np.maximum(0, v - 1), v)),
                            (2, 5, 11, np.clip(v, 5, 11)), (2, 6, 10, np.clip(v, 6, 10)),
                            (2, 4, 12, np.clip(v, 4, 12))]:
        out = unpack(oracle_mod.bs_apply(op, lo, hi, planes, mask))
        assert np.array_equal(out, ref), (op, lo, hi)


def test_topk_matches_sort(oracle_mod):
    rng = np.random.default_rng(1)
    for trial in range(200):
        ncol, incols = 512, 128
        scale = [3, 50, 4000, 300000][trial % 4]
        d = rng.integers(0, scale, ncol).astype(np.uint32)
        target = int(rng.integers(1, ncol - incols + 1))
        rec = np.sort(d[incols:])[::-1]
        kth = int(rec[target - 1])
        need = target - int((rec > kth).sum())
        for par in (True, False):
            th, nd = oracle_mod.topk(d, incols, target, par)
            assert th == max(kth, 1) and nd == need, (trial, par)


def test_topk_degenerate_paths(oracle_mod):
    d = np.zeros(64, np.uint32)
    d[8:] = np.arange(56, dtype=np.uint32)
    # fewer candidates than target: serial keeps need, parallel subtracts scan[1]
    th_s, nd_s = oracle_mod.topk(d, 8, 100, False)
    th_p, nd_p = oracle_mod.topk(d, 8, 100, True)
    assert th_s == th_p == 1
    assert nd_s == 100 and nd_p == (100 - 55) & 0xFFFFFFFF


def brute_kwta(v, k):
    rank = [sum(1 for j in range(32) if v[j] > v[l] or (v[j] == v[l] and j < l)) for l in range(32)]
    ball = sum(1 << l for l in range(32) if rank[l] < k)
    vk = max([v[l] if rank[l] == (k - 1) % 2**32 else 0 for l in range(32)])
    vk1 = max([v[l] if rank[l] == k else 0 for l in range(32)])
    return ball, vk, vk1


def test_kwta_brute_force(oracle_mod):
    rng = np.random.default_rng(2)
    for trial in range(300):
        hi = [2, 5, 100, 1 << 20][trial % 4]
        v = rng.integers(0, hi, 32).astype(np.uint32)
        k = int(rng.integers(0, 35))
        b, vk, vk1 = oracle_mod.kwta(v, k)
        assert (b, vk, vk1) == brute_kwta(v.tolist(), k)
        assert bin(b).count("1") == min(k, 32)


def test_parallel_oracle_is_bit_identical(oracle_mod):
    cfg = oracle_mod.parse(C0)
    a = oracle_mod.OracleNet(cfg, threads=1)
    b = oracle_mod.OracleNet(cfg, threads=4)
    for p in range(12):
        a.present(p, 1.0, 10, 5, True, 1000 + p)
        b.present(p, 1.0, 10, 5, True, 1000 + p)
        for name in ("HIST", "STATS", "BIAS", "THETA", "DRIVELOG"):
            assert np.array_equal(a.buf(name), b.buf(name)), (p, name)
    for i in range(4):
        assert np.array_equal(a.buf(f"W{i}"), b.buf(f"W{i}"))


def _pinned_cli(args):
    import oracle
    oracle.build()
    out = subprocess.run([oracle.CLI_PATH] + list(args) + ["threads=8"], capture_output=True, text=True,
                         check=True).stdout
    line = [l for l in out.splitlines() if l.startswith("PINNED")][0]
    return dict(kv.split("=") for kv in line.split()[1:]), out


@pytest.mark.slow
def test_untrained_control_full_scale():
    g = GOLD["untrained"]
    pin, _ = _pinned_cli(CONFIG_C + "stim=12 epochs=0 deep=1 wave=1 wavet=35 seed=0".split())
    assert pin["within"] == g["within_pct"]
    assert pin["drive6"] == g["drive_t6"]
    lo, hi = g["margin_range"]
    assert lo <= float(pin["margin"]) <= hi


@pytest.mark.slow
def test_golden_row_full_scale():
    g = GOLD["golden_row"]
    pin, _ = _pinned_cli(C1)
    assert pin["within"] == g["within_pct"]
    assert pin["recall"] == g["recall"]
    assert pin["other"] == g["best_other"]
    assert pin["sustained"] == g["sustained"]
    assert pin["drive6"] == g["drive_t6"]
