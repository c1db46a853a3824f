"""End-to-end parity of the HIP engine with the CPU oracle, presentation by
presentation: every per-step spike ballot, the stats and drive-log words,
theta, the intrinsic bias and all four synapse planes must be bit-identical.

Small networks (ncol=512) run the reference protocol's learning presentations
(mm:1009-1033) and a set of readout presentations that exercise every option
the readouts use (bias on/off, keep < 1, ktest, facil, stimulus override)."""
import numpy as np
import pytest

import axon_amd
import oracle
from conftest import C0, CONFIG_C

pytestmark = pytest.mark.gpu

SMALL = ["ncol=512", "incols=128"]
VARIANTS = {
    "c0_config_c": C0,
    "defaults_legacy_rule": SMALL + ["stimcols=64", "epochs=3"],                 # contrast=0, margingate=1
    "hetero_mode0_serialtopk": SMALL + CONFIG_C + ["hetero=1", "mode=0", "topkpar=0", "epochs=2", "ltdrep=2"],
    "mode2_biasevery": SMALL + CONFIG_C + ["mode=2", "dthr=3", "biasdownevery=3", "epochs=2", "seed=3"],
    "readout_knobs": SMALL + CONFIG_C + ["ktest=10", "facil=40", "epochs=2", "seed=1", "nand=3"],
    # the column gate at its boundaries: target = the 384 recurrent columns (gate by counting), and
    # beyond them (every workgroup runs the radix select)
    "target_all_recurrent": SMALL + CONFIG_C + ["target=384", "epochs=1", "seed=2"],
    "target_beyond_recurrent": SMALL + CONFIG_C + ["target=400", "epochs=1", "seed=2", "topkpar=0"],
    # accumulators past 2^16 (32 branches x (256 + 3000)): 32-bit accumulators instead of 16-bit
    "boost3000_acc32": SMALL + CONFIG_C + ["boost=3000", "epochs=1", "seed=6"],
}


def compare_present(g, o, T, tag, net=None):
    r = g
    assert np.array_equal(r.hist, o.buf("HIST")[: T * o.ncol].reshape(T, o.ncol)), f"{tag}: hist"
    assert np.array_equal(r.spikes, o.final_spikes(T)), f"{tag}: spikes"
    assert np.array_equal(r.stats[:1024], o.buf("STATS")[:1024]), f"{tag}: stats"
    assert np.array_equal(r.drive_log[:256], o.buf("DRIVELOG")[:256]), f"{tag}: driveLog"
    assert np.array_equal(r.theta, o.buf("THETA")[:2]), f"{tag}: theta"
    if net is not None:  # the last step's accumulators (16- or 32-bit on the device) and the stimulus
        assert np.array_equal(net.read("ACC"), o.buf("ACC")), f"{tag}: acc"
        assert np.array_equal(net.read("STIM"), o.buf("STIM")), f"{tag}: stim"


def compare_state(net, ora, tag, planes=True):
    for name in ["BIAS", "CONN"] + (["W0", "W1", "W2", "W3"] if planes else []):
        a, b = net.read(name), ora.buf(name)
        assert np.array_equal(a, b), f"{tag}: {name} differs at {np.flatnonzero(a != b)[:5]}"


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_protocol_parity(variant):
    args = VARIANTS[variant]
    cfg = axon_amd.Config.parse(args)
    ocfg = oracle.parse(args)
    print(f"parity {variant}: create", flush=True)
    net = axon_amd.Net(cfg)
    ora = oracle.OracleNet(ocfg, threads=4)
    compare_state(net, ora, f"{variant} init")
    T, clamp, tc, nstim = cfg.T, cfg.clampsteps, cfg.trainclamp, cfg.stim
    # settle (learn=false, freeze=false): no state change
    g = net.present(9 % nstim, 1.0, T, T, False, False, 999)
    ora.present(9 % nstim, 1.0, T, T, False, 999)
    compare_present(g, ora, T, f"{variant} settle")
    for ep in range(cfg.epochs):
        for p in range(nstim):
            g = net.present(p, 1.0, T, tc, True, False, 1000 + p)
            ora.present(p, 1.0, T, tc, True, 1000 + p)
            compare_present(g, ora, T, f"{variant} ep{ep} p{p}", net if p == 0 else None)
            compare_state(net, ora, f"{variant} ep{ep} p{p}", planes=(p == nstim - 1))
    g = net.present(0, 1.0, T, tc, True, False, 31337)
    ora.present(0, 1.0, T, tc, True, 31337)
    compare_present(g, ora, T, f"{variant} extra")
    compare_state(net, ora, f"{variant} extra")
    assert net.work_counters(reset=True) == ora.work_counters(reset=True), "work counters"
    print(f"parity {variant}: readouts", flush=True)
    refs = []
    for p in range(nstim):
        for keep, TT, cs, seed, ub in [(1.0, T, T, 2000 + p, 2), (0.75, T, clamp, 4400 + p, 0),
                                       (0.75, T, clamp, 4300 + p, 1), (1.0, min(35, cfg.max_t), clamp, 7000 + p, 2)]:
            g = net.present(p, keep, TT, cs, False, True, seed, ub)
            ora.present(p, keep, TT, cs, False, seed, ub)
            compare_present(g, ora, TT, f"{variant} readout p{p} seed{seed}", net if p == 0 else None)
            if seed == 2000 + p:
                refs.append(g.spikes.copy())
    # drive log with a reference ballot and multiplicities (deep / split blocks)
    mult = np.zeros(net.n, np.uint32)
    for r in refs:
        for c in range(cfg.incols, cfg.ncol):
            b = int(r[c])
            for l in range(32):
                if (b >> l) & 1:
                    mult[c * 32 + l] += 1
    net.set_mult(mult)
    ora.buf("MULT")[:] = mult
    for q in range(min(nstim, 4)):
        net.set_ref_ball(refs[q])
        ora.buf("REFBALL")[:] = refs[q]
        net.clear_drive_log()
        ora.buf("DRIVELOG")[:] = 0
        g = net.present(q, 1.0, 30, clamp, False, True, 8600 + q)
        ora.present(q, 1.0, 30, clamp, False, 8600 + q)
        compare_present(g, ora, 30, f"{variant} drivelog q{q}")
        assert np.array_equal(net.within(refs[q]), ora.within(refs[q]))
        pm = (np.arange(net.n) % 3 == 0).astype(np.uint32)
        assert np.array_equal(net.within(refs[q], pm), ora.within(refs[q], pm))
    # explicit stimulus override (extension of loadStim used by probe batches)
    stim = np.zeros(cfg.ncol, np.uint32)
    stim[np.random.default_rng(5).choice(cfg.incols, 40, replace=False)] = 0xFFFFFFFF
    g = net.present(0, 1.0, T, clamp, False, True, 0, 0, stim=stim)
    ora.present(0, 1.0, T, clamp, False, 0, 0, stim=stim)
    compare_present(g, ora, T, f"{variant} stim override")
    assert net.wstat() == ora.wstat()
    assert np.array_equal(net.indeg(), ora.indeg())
    assert np.array_equal(net.untouched(), ora.untouched())
    # the buffer checksum sees the accumulators as 32-bit values, whatever their width on the device
    assert axon_amd.checksum(net, "ACC") == int(axon_amd.row_digest(ora.buf("ACC")[None, :])[0])
    net.close()


@pytest.mark.parametrize("case", ["clamped", "unclamped_first_step", "stim_outside_inputs", "readout", "two_steps"])
def test_asynchronous_presentations_match_oracle(case):
    """Presentations without side statistics take the launch paths that fold
    timesteps together: a first timestep clamped to a stimulus inside the input
    columns runs inside the second timestep's propagation, each learning
    timestep's plasticity inside the next propagation and the last one's inside
    its emit.  An unclamped first timestep, or a stimulus with a bit outside the
    input columns, keeps the first timestep's own launch.  After every
    presentation the history rows, final spikes, accumulators, stimulus, bias
    and all four planes equal the oracle's."""
    args = C0[:]
    cfg = axon_amd.Config.parse(args)
    net = axon_amd.Net(cfg)
    ora = oracle.OracleNet(oracle.parse(args), threads=4)
    ncol, nin = cfg.ncol, cfg.incols
    T, clamp, learn = cfg.T, cfg.trainclamp, True
    if case == "unclamped_first_step":
        clamp = 0
    elif case == "readout":
        learn = False
    elif case == "two_steps":
        T, clamp = 2, 1
    for i in range(6):
        p = i % cfg.stim
        stim = None
        if case == "stim_outside_inputs":
            stim = np.zeros(ncol, np.uint32)
            stim[ora.stim_cols(p)] = 0xFFFFFFFF
            stim[nin + 7 + i] = 0x00F0F00F  # a recurrent column: its word is no spike vector of step 0
        print(f"async {case}: presentation {i}", flush=True)
        net.present(p, 1.0, T, clamp, learn, False, 1000 + i, stim=stim, outputs=False)
        ora.present(p, 1.0, T, clamp, learn, 1000 + i, stim=stim)
        net.sync()
        tag = f"async {case} presentation {i}"
        assert np.array_equal(net.read("HIST")[: T * ncol], ora.buf("HIST")[: T * ncol]), f"{tag}: hist"
        assert np.array_equal(net.read("SPK_A" if T % 2 == 0 else "SPK_B"), ora.final_spikes(T)), f"{tag}: spikes"
        assert np.array_equal(net.read("ACC"), ora.buf("ACC")), f"{tag}: acc"
        assert np.array_equal(net.read("STIM"), ora.buf("STIM")), f"{tag}: stim"
        compare_state(net, ora, tag)
    net.close()


def test_wdiag_parity():
    args = C0[:]
    net = axon_amd.Net(axon_amd.Config.parse(args))
    ora = oracle.OracleNet(oracle.parse(args), threads=4)
    for p in range(12):
        net.present(p, 1.0, 10, 5, True, False, 1000 + p, outputs=False)
        ora.present(p, 1.0, 10, 5, True, 1000 + p)
    net.sync()
    asm = np.zeros((12, 512), np.uint32)
    for p in range(12):
        asm[p] = net.present(p, 1.0, 10, 10, False, True, 2000 + p).spikes
        asm[p, :128] = 0
    memb = np.zeros(512 * 32, np.uint32)
    for c in range(128, 512):
        for l in range(32):
            m = 0
            for q in range(12):
                if (int(asm[q, c]) >> l) & 1:
                    m |= 1 << q
            memb[c * 32 + l] = m
    for sel in (0, 1, 2):
        assert np.array_equal(net.wdiag(memb, asm.ravel(), 12, sel), ora.wdiag(memb, asm.ravel(), 12, sel))
    # fewer assemblies than membership bits: a member of any assembly takes part, but its "within"
    # synapses come from assemblies below nstim only (mm:602, 618-622)
    assert np.array_equal(net.wdiag(memb, asm[:5].ravel(), 5, 0), ora.wdiag(memb, asm[:5].ravel(), 5, 0))
    # k_within's shared / exclusive split follows the multiplicity buffer
    mult = np.array([bin(int(m)).count("1") for m in memb], np.uint32)
    net.set_mult(mult)
    ora.buf("MULT")[:] = mult
    for q in (0, 7):
        assert np.array_equal(net.within(asm[q]), ora.within(asm[q]))
        assert np.array_equal(net.within(asm[q], (memb != 0).astype(np.uint32)),
                              ora.within(asm[q], (memb != 0).astype(np.uint32)))
    assert np.array_equal(net.indeg(), ora.indeg())
    assert np.array_equal(net.untouched(), ora.untouched())


def test_capacity_and_argument_errors():
    net = axon_amd.Net(axon_amd.Config.parse(C0))
    with pytest.raises(axon_amd.AxonError) as e:
        net.present(0, 1.0, 65, 5, False)
    assert e.value.status == 6  # AXON_ERR_CAPACITY, T beyond max_t
    with pytest.raises(axon_amd.AxonError) as e:
        net.present(12, 1.0, 10, 5, False)
    assert e.value.status == 5
    with pytest.raises(axon_amd.AxonError):
        net.present(0, 1.0, 10, 5, False, use_bias=3)
    big = axon_amd.Net(axon_amd.Config.parse(C0 + ["maxt=200"]))
    ora = oracle.OracleNet(oracle.parse(C0 + ["maxt=200"]), threads=4)
    r = big.present(0, 1.0, 150, 5, False, seed=7000)
    ora.present(0, 1.0, 150, 5, False, 7000)
    assert r.hist.shape == (150, 512)
    assert np.array_equal(r.hist.ravel(), ora.buf("HIST")[: 150 * 512])
    assert np.array_equal(r.stats, ora.buf("STATS"))
