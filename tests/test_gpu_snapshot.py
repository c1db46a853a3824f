"""Trained-network snapshots: a network saved and loaded again holds the same
planes, connectivity and bias, and every later presentation -- recall and
learning -- gives the same results as on the original; damaged files are
refused before any device work."""
import numpy as np
import pytest

import axon_amd
from conftest import C0

pytestmark = pytest.mark.gpu

STATE = ("W0", "W1", "W2", "W3", "CONN", "BIAS")


def trained():
    net = axon_amd.Net(axon_amd.Config.parse(C0 + ["epochs=2"]))
    c = net.config
    for ep in range(c.epochs):
        for p in range(c.stim):
            net.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p, outputs=False)
    net.sync()
    return net


def test_snapshot_holds_the_oracles_state(tmp_path):
    """A snapshot of the network after its training epochs, loaded back, holds
    the planes, connectivity, bias and stimulus columns of the CPU oracle's
    network after the same epochs; a recall on the loaded network is the
    oracle's."""
    import oracle

    args = C0 + ["epochs=2"]
    net = trained()
    ora = oracle.OracleNet(oracle.parse(args), threads=4)
    c = net.config
    for ep in range(c.epochs):
        for p in range(c.stim):
            ora.present(p, 1.0, c.T, c.trainclamp, True, 1000 + p)
    path = str(tmp_path / "net.axs")
    net.save(path)
    net.close()
    back = axon_amd.Net.load(path)
    for b in STATE:
        assert np.array_equal(back.read(b), ora.buf(b)), b
    for p in range(c.stim):
        assert np.array_equal(back.stim_cols(p), ora.stim_cols(p)), p
    r = back.present(3, 0.75, 35, 5, False, True, 7003, 0)
    ora.present(3, 0.75, 35, 5, False, 7003, 0)
    assert np.array_equal(r.hist.ravel(), ora.buf("HIST")[: 35 * c.ncol])
    back.close()


def test_save_load_round_trip(tmp_path):
    net = trained()
    path = str(tmp_path / "net.axs")
    net.save(path)
    print("snapshot bytes", (tmp_path / "net.axs").stat().st_size, flush=True)
    back = axon_amd.Net.load(path)
    assert back.config.as_dict() == net.config.as_dict()
    for b in STATE:
        assert axon_amd.checksum(back, b) == axon_amd.checksum(net, b), b
    cue = axon_amd.probe_cue(net, 3, 16, 77)
    r0 = net.present(0, 1.0, 35, 5, False, True, 0, 0, stim=cue)
    r1 = back.present(0, 1.0, 35, 5, False, True, 0, 0, stim=cue)
    assert np.array_equal(r0.hist, r1.hist) and np.array_equal(r0.spikes, r1.spikes)
    c = net.config
    for p in (0, 5, 11):  # learning continues identically
        a = net.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p)
        b = back.present(p, 1.0, c.T, c.trainclamp, True, False, 1000 + p)
        assert np.array_equal(a.hist, b.hist), f"pattern {p}"
    for b in STATE:
        assert axon_amd.checksum(back, b) == axon_amd.checksum(net, b), f"{b} after learning"
    back.close()
    net.close()


def test_damaged_snapshots_are_refused(tmp_path):
    net = trained()
    path = tmp_path / "net.axs"
    net.save(str(path))
    with pytest.raises(axon_amd.AxonError) as e:
        net.save(str(tmp_path / "missing-dir" / "x.axs"))
    assert e.value.status == 8  # AXON_ERR_IO
    net.close()
    data = bytearray(path.read_bytes())
    cases = {}
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x40  # a plane byte
    cases["flipped"] = flipped
    cases["truncated"] = data[: len(data) - 100]
    cases["trailing"] = data + b"x"
    version = bytearray(data)
    version[8] = 2  # version field follows the 8-byte magic
    cases["version"] = version
    for name, blob in cases.items():
        bad = tmp_path / f"{name}.axs"
        bad.write_bytes(bytes(blob))
        with pytest.raises(axon_amd.AxonError) as e:
            axon_amd.Net.load(str(bad))
        assert e.value.status == 9, name  # AXON_ERR_BAD_SNAPSHOT
