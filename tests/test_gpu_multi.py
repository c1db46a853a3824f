"""Many networks trained in shared launches must each end in exactly the state
of training them alone (planes, bias, work counters), and that state must match
the CPU oracle.  Networks differ in seed, number of stored patterns and epochs,
so their work queues have different lengths."""
import numpy as np
import pytest

import axon_amd
import oracle
from conftest import CONFIG_C

pytestmark = pytest.mark.gpu

SMALL = ["ncol=512", "incols=128"]
SPECS = [(4, 0, 3), (8, 1, 2), (12, 2, 1), (6, 3, 4), (10, 4, 2), (5, 5, 0)]  # (stim, seed, epochs)


def cfg_args(stim, seed):
    return SMALL + CONFIG_C + [f"stim={stim}", f"seed={seed}"]


def train_alone(net, epochs):
    c = net.config
    for ep in range(epochs):
        for p in range(c.stim):
            net.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p, outputs=False)
    net.present(0, 1.0, c.T, c.trainclamp, True, False, 31337, outputs=False)
    net.sync()


def test_multi_matches_plain_and_oracle():
    batched = [axon_amd.Net(axon_amd.Config.parse(cfg_args(s, sd))) for s, sd, _ in SPECS]
    mb = axon_amd.MultiNet(batched)
    ms = mb.train([e for _, _, e in SPECS], extra=True)
    print(f"multi: {len(SPECS)} networks trained in {ms:.1f} ms", flush=True)
    for (stim, seed, ep), bn in zip(SPECS, batched):
        alone = axon_amd.Net(axon_amd.Config.parse(cfg_args(stim, seed)))
        train_alone(alone, ep)
        for name in ("W0", "W1", "W2", "W3", "BIAS"):
            a, b = bn.read(name), alone.read(name)
            assert np.array_equal(a, b), f"stim={stim} seed={seed}: {name}"
            assert axon_amd.checksum(bn, name) == axon_amd.checksum(alone, name)
        assert bn.work_counters() == alone.work_counters()
        alone.close()
    for (stim, seed, ep), bn in list(zip(SPECS, batched))[:2]:
        ora = oracle.OracleNet(oracle.parse(cfg_args(stim, seed)), threads=4)
        for e in range(ep):
            for p in range(stim):
                ora.present(p, 1.0, 10, 5, True, 1000 + p)
        ora.present(0, 1.0, 10, 5, True, 31337)
        for name in ("W0", "W1", "W2", "W3", "BIAS"):
            assert np.array_equal(bn.read(name), ora.buf(name)), f"oracle {stim}/{seed}: {name}"
    # the batch can be trained again (a second schedule continues from the current state)
    mb.train([1] * len(SPECS), extra=False)
    mb.close()


def test_multi_rejects_mismatched_dynamics():
    a = axon_amd.Net(axon_amd.Config.parse(cfg_args(4, 0)))
    b = axon_amd.Net(axon_amd.Config.parse(cfg_args(4, 1) + ["k=20"]))
    with pytest.raises(axon_amd.AxonError):
        axon_amd.MultiNet([a, b])
    with pytest.raises(axon_amd.AxonError):
        axon_amd.MultiNet([a, a])


@pytest.mark.parametrize("mode_args", [["mode=0", "hetero=1"], ["mode=1", "dthr=2", "boost=3"], ["mode=2", "dthr=3"],
                                       ["target=384"], ["target=400", "topkpar=0"], ["ncol=520"],
                                       ["trainclamp=0"], ["T=2", "trainclamp=1"],
                                       ["T=3", "trainclamp=2", "contrast=0"]],
                         ids=["mode0-hetero", "mode1-dthr2", "mode2", "gate-count-all-recurrent",
                              "gate-topk-beyond-recurrent", "gate-topk-ncol520", "unclamped-first-step",
                              "two-steps", "three-steps-legacy-rule"])
def test_multi_dendritic_modes(mode_args):
    """Every branch-output mode of the two-column propagation, the per-network top-K launch at the
    target boundaries (every recurrent column, beyond them) and at a column count that is not a multiple
    of 32, and the round's launch structure: a first timestep that is not clamped (its own launch), two
    timesteps (the first inside the second's propagation, whose emit also runs the plasticity) and an odd
    timestep count under the legacy rule (every timestep learns): batched training equals training
    alone."""
    specs = [(4, 0, 2), (6, 1, 1), (5, 2, 3)]
    batched = [axon_amd.Net(axon_amd.Config.parse(cfg_args(s, sd) + mode_args)) for s, sd, _ in specs]
    mb = axon_amd.MultiNet(batched)
    mb.train([e for _, _, e in specs], extra=True)
    print(f"multi {mode_args}: trained", flush=True)
    for (stim, seed, ep), bn in zip(specs, batched):
        alone = axon_amd.Net(axon_amd.Config.parse(cfg_args(stim, seed) + mode_args))
        train_alone(alone, ep)
        for name in ("W0", "W1", "W2", "W3", "BIAS"):
            assert np.array_equal(bn.read(name), alone.read(name)), f"{mode_args} stim={stim} seed={seed}: {name}"
        assert bn.work_counters() == alone.work_counters()
        alone.close()
    mb.close()
