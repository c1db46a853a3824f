"""The allocation-free probe run (axon_probes_run_noalloc) against the entry
point it mirrors (axon_probes_run): same network, same cues, identical spike
histories and final ballots, on two calls in a row (the second smaller, so
state the first call left in the shared static buffer must not leak into it),
plus the capacity check.  Device buffers for the inputs and outputs are
allocated by the test itself through the HIP runtime the library is linked
against (torch bundles a runtime of its own, which must not share the process
with it); the library allocates nothing for the run."""
import ctypes

import numpy as np
import pytest

import axon_amd
from conftest import C0, CONFIG_C

pytestmark = pytest.mark.gpu


class DevBuf:
    """hipMalloc'd words, filled from / read back to numpy with hipMemcpy."""

    _hip = None

    def __init__(self, words: int, data=None):
        if DevBuf._hip is None:
            axon_amd.lib()  # loads libamdhip64.so.7; the same soname returns that copy
            DevBuf._hip = ctypes.CDLL("libamdhip64.so.7")
        self.words = words
        self.ptr = ctypes.c_void_p()
        assert DevBuf._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(words, 1) * 4)) == 0
        if data is not None:
            a = np.ascontiguousarray(data, np.uint32).ravel()
            assert a.size == words
            assert DevBuf._hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_size_t(words * 4), 1) == 0  # host to device

    def numpy(self) -> np.ndarray:
        a = np.zeros(self.words, np.uint32)
        assert DevBuf._hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                     ctypes.c_size_t(self.words * 4), 2) == 0  # device to host
        return a

    def free(self):
        if self.ptr:
            DevBuf._hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _trained(args, epochs):
    net = axon_amd.Net(axon_amd.Config.parse(args))
    cfg = net.config
    for _ in range(epochs):
        for p in range(cfg.stim):
            net.present(p, 1.0, cfg.T, cfg.trainclamp, True, False, 1000 + p, outputs=False)
    net.sync()
    return net


def _cues(net, n, seed0):
    cfg = net.config
    out = []
    for i in range(n):
        if i % 5 == 4:
            out.append(axon_amd.probe_cue(net, -1, 0, seed0 + i))
        else:
            out.append(axon_amd.probe_cue(net, i % cfg.stim, [0, 6, 16, 32][i % 4], seed0 + i))
    return np.stack(out)


def _noalloc(net, stims, T, clamp, tile, with_hist=True):
    count, ncol = stims.shape
    d_stims = DevBuf(count * ncol, stims)  # hipMemcpy is synchronous: uploaded before the run reads it
    d_hist = DevBuf(count * T * ncol) if with_hist else None
    d_spk = DevBuf(count * ncol)
    axon_amd.probes_run_noalloc(net, count, d_stims.ptr.value, T, clamp, 0, tile,
                                d_hist.ptr.value if with_hist else 0, d_spk.ptr.value)
    net.sync()
    spk = d_spk.numpy().reshape(count, ncol)
    hist = d_hist.numpy().reshape(count, T, ncol) if with_hist else None
    for b in (d_stims, d_hist, d_spk):
        if b is not None:
            b.free()
    return spk, hist


def _reference(net, stims, T, clamp, tile):
    pb = axon_amd.Probes(net, slots=stims.shape[0], max_t=T, tile=tile)
    spk, hist, _ = pb.run(stims, T, clamp, use_bias=0)
    pb.close()
    return spk, hist


@pytest.mark.parametrize("tile", [4, 8, 32])
def test_noalloc_matches_probe_batch_twice(tile):
    net = _trained(C0, 2)
    for call, (count, T) in enumerate([(40, 35), (9, 12)]):  # second call smaller
        stims = _cues(net, count, 500 + 100 * call)
        print(f"noalloc tile {tile} call {call}: {count} probes x {T} steps", flush=True)
        spk, hist = _noalloc(net, stims, T, 5, tile)
        rspk, rhist = _reference(net, stims, T, 5, tile)
        assert np.array_equal(hist, rhist), f"tile {tile} call {call}: histories differ"
        assert np.array_equal(spk, rspk), f"tile {tile} call {call}: final ballots differ"
    net.close()


def test_noalloc_without_history_output_full_scale():
    net = axon_amd.Net(axon_amd.Config.parse(CONFIG_C + ["stim=12", "seed=0"]))
    for call, (count, T, tile) in enumerate([(16, 20, 8), (5, 7, 4)]):
        stims = _cues(net, count, 40 + call)
        print(f"noalloc full scale call {call}: {count} probes x {T} steps, tile {tile}", flush=True)
        spk, _ = _noalloc(net, stims, T, 5, tile, with_hist=False)
        rspk, _ = _reference(net, stims, T, 5, tile)
        assert np.array_equal(spk, rspk), f"full scale call {call}: final ballots differ"
    net.close()


def noalloc_state_words(ncol: int, count: int, tile: int, T: int, with_hist: bool) -> int:
    """Words of static-buffer state a run needs: per slot (count rounded up to
    the tile) spikes, drives, accumulators, spiking-column bits, theta and five
    step counters, each piece rounded up to 64 words; the history too when the
    caller passes no history buffer."""
    ns = -(-count // tile) * tile
    al = lambda w: (w + 63) // 64 * 64  # noqa: E731
    pieces = [ns * ncol, ns * (-(-ncol // 32)), ns * ncol * 32, ns * ncol, 2 * ns] + [ns] * 5
    if not with_hist:
        pieces.append(ns * T * ncol)
    return sum(al(w) for w in pieces)


def test_noalloc_rejects_state_beyond_the_static_buffer():
    net = axon_amd.Net(axon_amd.Config.parse(C0))
    ncol, T, tile = 512, 10, 8
    cap = axon_amd.NOALLOC_SCRATCH_BYTES // 4
    fits = max(c for c in range(1, 4096) if noalloc_state_words(ncol, c, tile, T, True) <= cap)
    count = 2 * fits  # far beyond the buffer
    assert noalloc_state_words(ncol, count, tile, T, True) > 1.9 * cap
    # real buffers of the sizes the call names, so every pointer passed is valid
    d_stims = DevBuf(count * ncol, np.zeros(count * ncol, np.uint32))
    d_hist = DevBuf(count * T * ncol)
    d_spk = DevBuf(count * ncol)
    with pytest.raises(axon_amd.AxonError) as e:
        axon_amd.probes_run_noalloc(net, count, d_stims.ptr.value, T, 5, 0, tile, d_hist.ptr.value, d_spk.ptr.value)
    assert e.value.status == 6  # AXON_ERR_CAPACITY
    for b in (d_stims, d_hist, d_spk):
        b.free()
    net.close()
