"""axon_amd -- Python mirror of the MI355X AXON engine's C ABI (libaxon.so).

The engine itself is C++/HIP (``axon_amd/csrc``); this module only binds the
C ABI declared in ``include/axon_api.h`` with ctypes, so tests and the bench
drive exactly the entry points a C or Objective-C++ caller would.  There is no
fallback: if ``libaxon.so`` is missing or no gfx950 device is usable, calls
raise ``AxonError``.

The reference's interface for this path is its key=value CLI and the
``present(p, keep, T, clampSteps, learn, freeze, seed, useBias)`` routine of
src/axon.mm (mm:660-667, mm:815-888); ``Config.parse`` and ``Net.present``
keep those argument names and meanings.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaxon.so")
CLI_PATH = os.path.join(HERE, "bin", "axon")

CONFIG_FIELDS = (
    "ncol incols nand target k dthr boost mode "
    "epochs T clampsteps stim seed quiet split unionchar "
    "biasup biasdown biasshift biasdownevery inbr stimcols diag profile "
    "topkpar vtopk contrast bandlo bandhi margingate hetero vbs "
    "ltdrep ltdevery ktest facil deep wave wavet trainclamp max_t"
).split()

# Configuration C of the paper (README.md:45-47 of the reference).
CONFIG_C = (
    "nand=4 target=64 k=24 dthr=1 mode=1 inbr=4 stimcols=64 contrast=1 margingate=0 "
    "ltdrep=1 ltdevery=2 biasup=73 biasdown=1 T=10 clampsteps=5"
).split()


class AxonError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"axon status {status}: {message}")
        self.status = status


class ConfigStruct(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint32) for f in CONFIG_FIELDS] + [("reserved", ctypes.c_uint32 * 7)]


class PresentArgs(ctypes.Structure):
    _fields_ = [
        ("pattern", ctypes.c_uint32), ("T", ctypes.c_uint32), ("clamp_steps", ctypes.c_uint32),
        ("learn", ctypes.c_uint32), ("freeze", ctypes.c_uint32), ("seed", ctypes.c_uint32),
        ("use_bias", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("keep", ctypes.c_double),
        ("stim", ctypes.POINTER(ctypes.c_uint32)),
    ]


U32P = ctypes.POINTER(ctypes.c_uint32)


class PresentOut(ctypes.Structure):
    _fields_ = [
        ("spikes", U32P), ("hist", U32P), ("stats", U32P), ("drive_log", U32P), ("theta", U32P),
        ("ms_per_step", ctypes.c_double),
    ]


BUFFERS = ("W0 W1 W2 W3 CONN SPK_A SPK_B DRIVE BBITS ACC BIAS MARG REFBALL DRIVELOG MULT "
           "STIM STATS HIST THETA").split()
PRESENT_ASYNC = 1
PROBE_TILES = (1, 2, 4, 8, 16, 32, 64)  # axon_probe_autotune's order (AXON_PROBE_TILES)
PROBE_WAVES = (0, 32, 16)  # waves per CU, its order of occupancies (AXON_PROBE_OCCUPANCIES)

_lib = None


def lib() -> ctypes.CDLL:
    """Load libaxon.so (built by __graft_entry__.build / axon_amd/Makefile)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AxonError(-1, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.axon_status_str.restype = ctypes.c_char_p
        L.axon_last_error.restype = ctypes.c_char_p
        L.axon_abi_version.restype = ctypes.c_uint32
        L.axon_config_defaults.restype = None
        L.axon_net_destroy.restype = None
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        L = lib()
        raise AxonError(status, f"{L.axon_status_str(status).decode()}: {L.axon_last_error().decode()}")


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U32P)


@dataclass
class Config:
    """All key=value tunables of the reference CLI (mm:687-735)."""

    raw: ConfigStruct

    @classmethod
    def parse(cls, args: Sequence[str]) -> "Config":
        """argu() semantics: prefix match + '=', last occurrence wins (mm:660-667)."""
        argv = ["axon"] + list(args)
        arr = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
        c = ConfigStruct()
        _check(lib().axon_config_parse(len(argv), arr, ctypes.byref(c)))
        return cls(c)

    @classmethod
    def defaults(cls) -> "Config":
        c = ConfigStruct()
        lib().axon_config_defaults(ctypes.byref(c))
        return cls(c)

    def validate(self) -> None:
        _check(lib().axon_config_validate(ctypes.byref(self.raw)))

    def __getattr__(self, name):
        if name in CONFIG_FIELDS:
            return getattr(self.raw, name)
        raise AttributeError(name)

    def as_dict(self) -> dict:
        return {f: getattr(self.raw, f) for f in CONFIG_FIELDS}


@dataclass
class PresentResult:
    spikes: np.ndarray
    hist: np.ndarray
    stats: np.ndarray
    drive_log: np.ndarray
    theta: np.ndarray
    ms_per_step: float


class Net:
    """One network resident in HBM (planes, connectivity, bias, stimuli)."""

    def __init__(self, config: Config, _handle: Optional[ctypes.c_void_p] = None):
        self.config = config
        if _handle is None:
            self._h = ctypes.c_void_p()
            _check(lib().axon_net_create(ctypes.byref(config.raw), ctypes.byref(self._h)))
        else:
            self._h = _handle
        self.ncol = config.ncol
        self.n = config.ncol * 32

    def save(self, path: str) -> None:
        """Write a versioned snapshot (planes, conn, bias, stimulus columns; axon_net_save)."""
        _check(lib().axon_net_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path: str) -> "Net":
        """A network created from a snapshot file (axon_net_load)."""
        h = ctypes.c_void_p()
        _check(lib().axon_net_load(os.fsencode(path), ctypes.byref(h)))
        c = ConfigStruct()
        st = lib().axon_net_config(h, ctypes.byref(c))
        if st != 0:
            lib().axon_net_destroy(h)
            _check(st)
        return cls(Config(c), _handle=h)

    def close(self) -> None:
        if self._h:
            lib().axon_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def words(self, which: str) -> int:
        w = ctypes.c_size_t()
        _check(lib().axon_buffer_words(self._h, BUFFERS.index(which), ctypes.byref(w)))
        return w.value

    def read(self, which: str, nwords: Optional[int] = None) -> np.ndarray:
        n = self.words(which) if nwords is None else nwords
        out = np.empty(n, dtype=np.uint32)
        _check(lib().axon_read_buffer(self._h, BUFFERS.index(which), out.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_size_t(n * 4)))
        return out

    def write(self, which: str, data: np.ndarray) -> None:
        data = np.ascontiguousarray(data, dtype=np.uint32)
        _check(lib().axon_write_buffer(self._h, BUFFERS.index(which), data.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.c_size_t(data.size * 4)))

    def stim_cols(self, p: int) -> np.ndarray:
        out = np.empty(self.config.stimcols, dtype=np.uint32)
        _check(lib().axon_stim_cols(self._h, ctypes.c_uint32(p), _ptr(out)))
        return out

    def present(self, p: int, keep: float, T: int, clamp_steps: int, learn: bool, freeze: bool = True,
                seed: int = 0, use_bias: int = 2, stim: Optional[np.ndarray] = None,
                outputs: bool = True) -> Optional[PresentResult]:
        """present() of mm:815-888.  outputs=False enqueues without waiting."""
        a = PresentArgs(p, T, clamp_steps, int(bool(learn)), int(bool(freeze)), seed, use_bias,
                        0 if outputs else PRESENT_ASYNC, float(keep), None)
        if stim is not None:
            stim = np.ascontiguousarray(stim, dtype=np.uint32)
            a.stim = _ptr(stim)
        o = PresentOut()
        res = None
        if outputs:
            res = PresentResult(np.zeros(self.ncol, np.uint32), np.zeros(max(T, 1) * self.ncol, np.uint32),
                                np.zeros(self.words("STATS"), np.uint32), np.zeros(self.words("DRIVELOG"), np.uint32),
                                np.zeros(2, np.uint32), 0.0)
            o.spikes, o.hist, o.stats = _ptr(res.spikes), _ptr(res.hist), _ptr(res.stats)
            o.drive_log, o.theta = _ptr(res.drive_log), _ptr(res.theta)
        _check(lib().axon_present(self._h, ctypes.byref(a), ctypes.byref(o)))
        if res is not None:
            res.ms_per_step = o.ms_per_step
            res.hist = res.hist[: T * self.ncol].reshape(T, self.ncol)
        return res

    def work_counters(self, reset: bool = False):
        """(tiles read by propagation, tiles rewritten by plasticity); 8192 synapses per tile."""
        out = (ctypes.c_uint64 * 2)()
        _check(lib().axon_work_counters(self._h, out, ctypes.c_uint32(int(reset))))
        return int(out[0]), int(out[1])

    def line_counters(self, reset: bool = False):
        """(128-byte W3 lines read by propagation, word lines rewritten by plasticity,
        each 4 planes x 128 bytes read and written).  reset also clears work_counters."""
        out = (ctypes.c_uint64 * 2)()
        _check(lib().axon_line_counters(self._h, out, ctypes.c_uint32(int(reset))))
        return int(out[0]), int(out[1])

    def last_step_counters(self):
        """(tiles, word lines) of the plasticity counts done at the last timestep of
        presentations without side statistics, inside that timestep's emit launch (the
        other timesteps' plasticity runs inside the next propagation launch)."""
        out = (ctypes.c_uint64 * 2)()
        _check(lib().axon_last_step_counters(self._h, out))
        return int(out[0]), int(out[1])

    def set_profiling(self, enable: bool) -> None:
        _check(lib().axon_net_set_profiling(self._h, ctypes.c_uint32(int(enable))))

    def kernel_times(self, reset: bool = False):
        """Summed device ms and launch counts of (propagation, first-step launch, top-K + emit,
        plasticity); without side statistics the last slot is the last timestep's emit launch,
        which runs that timestep's plasticity too (its emit is not in the third)."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_uint64 * 4)()
        _check(lib().axon_kernel_times(self._h, ms, n, ctypes.c_uint32(int(reset))))
        return list(ms), [int(x) for x in n]

    def sync(self) -> None:
        _check(lib().axon_sync(self._h))

    def set_ref_ball(self, ball: np.ndarray) -> None:
        _check(lib().axon_set_ref_ball(self._h, _ptr(np.ascontiguousarray(ball, np.uint32))))

    def set_mult(self, mult: np.ndarray) -> None:
        _check(lib().axon_set_mult(self._h, _ptr(np.ascontiguousarray(mult, np.uint32))))

    def clear_drive_log(self) -> None:
        _check(lib().axon_clear_drive_log(self._h))

    def set_topk_parallel(self, par: bool) -> None:
        _check(lib().axon_net_set_topk_parallel(self._h, ctypes.c_uint32(int(par))))

    def wstat(self) -> int:
        v = ctypes.c_uint32()
        _check(lib().axon_wstat(self._h, ctypes.byref(v)))
        return v.value

    def within(self, ball: np.ndarray, post_mask: Optional[np.ndarray] = None) -> np.ndarray:
        out = np.zeros(16, np.uint32)
        ball = np.ascontiguousarray(ball, np.uint32)
        pm = None if post_mask is None else np.ascontiguousarray(post_mask, np.uint32)
        _check(lib().axon_within(self._h, _ptr(ball), _ptr(pm), _ptr(out)))
        return out

    def indeg(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint32)
        _check(lib().axon_indeg(self._h, _ptr(out)))
        return out

    def untouched(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint32)
        _check(lib().axon_untouched(self._h, _ptr(out)))
        return out

    def wdiag(self, memb: np.ndarray, asm_spk: np.ndarray, nstim: int, sel: int) -> np.ndarray:
        out = np.zeros(48, np.uint32)
        _check(lib().axon_wdiag(self._h, _ptr(np.ascontiguousarray(memb, np.uint32)),
                                _ptr(np.ascontiguousarray(asm_spk, np.uint32)), ctypes.c_uint32(nstim),
                                ctypes.c_uint32(sel), _ptr(out)))
        return out


class Probes:
    """A probe batch: `slots` readout presentations against one resident network."""

    def __init__(self, net: Net, slots: int, max_t: int, tile: int = 4, waves_per_cu: int = 0):
        self.net = net
        self._h = ctypes.c_void_p()
        _check(lib().axon_probes_create(net._h, ctypes.c_uint32(slots), ctypes.c_uint32(max_t),
                                        ctypes.c_uint32(tile), ctypes.byref(self._h)))
        _check(lib().axon_probes_set_occupancy(self._h, ctypes.c_uint32(waves_per_cu)))
        s, m, t = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().axon_probes_info(self._h, ctypes.byref(s), ctypes.byref(m), ctypes.byref(t)))
        self.slots, self.max_t, self.tile = s.value, m.value, t.value

    def close(self) -> None:
        if self._h:
            lib().axon_probes_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, first: int, stims: np.ndarray, T, clamp_steps: int, use_bias: int = 0) -> None:
        stims = np.ascontiguousarray(stims, np.uint32).reshape(-1, self.net.ncol)
        count = stims.shape[0]
        Ts = np.ascontiguousarray(np.broadcast_to(np.asarray(T, np.uint32), (count,)), np.uint32)
        _check(lib().axon_probes_load(self._h, ctypes.c_uint32(first), ctypes.c_uint32(count), _ptr(stims),
                                      _ptr(Ts), ctypes.c_uint32(clamp_steps), ctypes.c_uint32(use_bias)))

    def step(self, steps: int) -> None:
        _check(lib().axon_probes_step(self._h, ctypes.c_uint32(steps)))

    def sync(self) -> None:
        _check(lib().axon_probes_sync(self._h))

    def read(self, first: int, count: int, hist: bool = True):
        spk = np.zeros((count, self.net.ncol), np.uint32)
        h = np.zeros((count, self.max_t, self.net.ncol), np.uint32) if hist else None
        done = np.zeros(count, np.uint32)
        _check(lib().axon_probes_read(self._h, ctypes.c_uint32(first), ctypes.c_uint32(count), _ptr(spk),
                                      _ptr(h) if hist else None, _ptr(done)))
        return spk, h, done

    def run(self, stims: np.ndarray, T: int, clamp_steps: int, use_bias: int = 0, hist: bool = True):
        stims = np.ascontiguousarray(stims, np.uint32).reshape(-1, self.net.ncol)
        count = stims.shape[0]
        h = np.zeros((count, T, self.net.ncol), np.uint32) if hist else None
        spk = np.zeros((count, self.net.ncol), np.uint32)
        ms = ctypes.c_double()
        _check(lib().axon_probes_run(self._h, ctypes.c_uint32(count), _ptr(stims), ctypes.c_uint32(T),
                                     ctypes.c_uint32(clamp_steps), ctypes.c_uint32(use_bias),
                                     _ptr(h) if hist else None, _ptr(spk), ctypes.byref(ms)))
        return spk, h, ms.value


NOALLOC_SCRATCH_BYTES = 32 << 20  # AXON_NOALLOC_SCRATCH_BYTES


def probes_run_noalloc(net: Net, count: int, d_stims: int, T: int, clamp_steps: int, use_bias: int, tile: int,
                       d_hist: int = 0, d_spikes: int = 0, stream: int = 0) -> None:
    """axon_probes_run_noalloc: device addresses in, nothing allocated; enqueues only."""
    _check(lib().axon_probes_run_noalloc(
        net._h, ctypes.c_uint32(count), ctypes.c_void_p(d_stims), ctypes.c_uint32(T), ctypes.c_uint32(clamp_steps),
        ctypes.c_uint32(use_bias), ctypes.c_uint32(tile), ctypes.c_void_p(d_hist or None),
        ctypes.c_void_p(d_spikes or None), ctypes.c_void_p(stream or None)))


class MultiNet:
    """Independent networks trained in shared launches (per-network work queues)."""

    def __init__(self, nets: Sequence[Net]):
        self.nets = list(nets)
        arr = (ctypes.c_void_p * len(self.nets))(*[n._h.value for n in self.nets])
        self._h = ctypes.c_void_p()
        _check(lib().axon_multi_create(arr, ctypes.c_uint32(len(self.nets)), ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().axon_multi_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, epochs: Sequence[int], extra: bool = True) -> float:
        ep = np.ascontiguousarray(epochs, np.uint32)
        ms = ctypes.c_double()
        _check(lib().axon_multi_train(self._h, _ptr(ep), ctypes.c_uint32(int(extra)), ctypes.byref(ms)))
        return ms.value


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots_per_net", ctypes.c_uint32), ("max_t", ctypes.c_uint32), ("tile", ctypes.c_uint32),
                ("steps_per_tick", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
                ("waves_per_cu", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class Server:
    """Continuous-batching recall server over resident networks."""

    def __init__(self, nets: Sequence[Net], slots_per_net: int = 64, max_t: int = 200, tile: int = 4,
                 steps_per_tick: int = 5, hbm_budget: int = 0, waves_per_cu: int = 0):
        self.nets = list(nets)
        self.ncol = self.nets[0].ncol
        arr = (ctypes.c_void_p * len(self.nets))(*[n._h.value for n in self.nets])
        cfg = ServerConfig(slots_per_net, max_t, tile, steps_per_tick, hbm_budget, waves_per_cu)
        self._h = ctypes.c_void_p()
        L = lib()
        L.axon_server_clock.restype = ctypes.c_double
        L.axon_server_clock.argtypes = [ctypes.c_void_p]
        _check(L.axon_server_create(arr, ctypes.c_uint32(len(self.nets)), ctypes.byref(cfg), ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().axon_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return lib().axon_server_clock(self._h)

    def submit(self, net: int, stim: np.ndarray, T: int, clamp_steps: int = 5, use_bias: int = 0,
               arrival: float = -1.0) -> int:
        stim = np.ascontiguousarray(stim, np.uint32)
        rid = ctypes.c_uint64()
        _check(lib().axon_server_submit(self._h, ctypes.c_uint32(net), _ptr(stim), ctypes.c_uint32(T),
                                        ctypes.c_uint32(clamp_steps), ctypes.c_uint32(use_bias),
                                        ctypes.c_double(arrival), ctypes.byref(rid)))
        return rid.value

    def tick(self) -> int:
        n = ctypes.c_uint32()
        _check(lib().axon_server_tick(self._h, ctypes.byref(n)))
        return n.value

    def completed(self) -> list:
        """Ids of the queries the last tick finished."""
        n = ctypes.c_uint32()
        _check(lib().axon_server_completed(self._h, None, 0, ctypes.byref(n)))
        ids = np.zeros(max(n.value, 1), np.uint64)
        _check(lib().axon_server_completed(self._h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                           ctypes.c_uint32(n.value), ctypes.byref(n)))
        return [int(x) for x in ids[: n.value]]

    def result(self, rid: int, max_t: int):
        """(spikes, row digests, latency seconds, T) or None while pending."""
        spk = np.zeros(self.ncol, np.uint32)
        dig = np.zeros(max_t, np.uint64)
        lat, T = ctypes.c_double(), ctypes.c_uint32()
        st = lib().axon_server_result(self._h, ctypes.c_uint64(rid), _ptr(spk),
                                      dig.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(lat),
                                      ctypes.byref(T))
        if st == 7:  # AXON_ERR_NOT_READY
            return None
        _check(st)
        return spk, dig[: T.value], lat.value, T.value

    def config(self) -> dict:
        """The configuration in use (tile and occupancy as chosen for tile=0)."""
        c = ServerConfig()
        _check(lib().axon_server_get_config(self._h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_ if k != "reserved"}

    def status(self):
        q, f, r = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().axon_server_status(self._h, ctypes.byref(q), ctypes.byref(f), ctypes.byref(r)))
        return q.value, f.value, r.value


def _lowbias32(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def row_digest(rows: np.ndarray) -> np.ndarray:
    """The engine's order-independent 64-bit checksum of each row of u32 words."""
    rows = np.atleast_2d(np.asarray(rows, np.uint32)).astype(np.uint64)
    i = np.arange(rows.shape[1], dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = rows * (2 * i + 1) + _lowbias32(rows ^ i)
        return v.sum(axis=1, dtype=np.uint64)


def checksum(net: Net, which: str) -> int:
    v = ctypes.c_uint64()
    _check(lib().axon_checksum(net._h, BUFFERS.index(which), ctypes.byref(v)))
    return v.value


def probe_autotune(net: Net, stims: np.ndarray, T: int, clamp_steps: int, use_bias: int = 0):
    """Fastest (tile, waves per CU) for this batch and the time of every pair:
    (tile, waves, {(tile, waves): ms})."""
    stims = np.ascontiguousarray(stims, np.uint32).reshape(-1, net.ncol)
    best, waves = ctypes.c_uint32(), ctypes.c_uint32()
    ms = (ctypes.c_double * (len(PROBE_TILES) * len(PROBE_WAVES)))()
    _check(lib().axon_probe_autotune(net._h, ctypes.c_uint32(stims.shape[0]), _ptr(stims), ctypes.c_uint32(T),
                                     ctypes.c_uint32(clamp_steps), ctypes.c_uint32(use_bias), ctypes.byref(best),
                                     ctypes.byref(waves), ms))
    table = {(t, w): ms[i * len(PROBE_WAVES) + j] for i, t in enumerate(PROBE_TILES) for j, w in enumerate(PROBE_WAVES)}
    return best.value, waves.value, table


class TuneEntry(ctypes.Structure):
    _fields_ = [("ncol", ctypes.c_uint32), ("probes", ctypes.c_uint32), ("T", ctypes.c_uint32),
                ("tile", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32), ("reps", ctypes.c_uint32),
                ("ms", ctypes.c_double), ("spread_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def probe_autotune_sizes(net: Net, stims: np.ndarray, T: int, clamp_steps: int, sizes: Sequence[int],
                         use_bias: int = 0) -> list:
    """The autotuner for the first `size` stimuli, per size: one tuning-table
    entry each (as dicts)."""
    stims = np.ascontiguousarray(stims, np.uint32).reshape(-1, net.ncol)
    sz = np.ascontiguousarray(sizes, np.uint32)
    out = (TuneEntry * len(sz))()
    _check(lib().axon_probe_autotune_sizes(net._h, ctypes.c_uint32(stims.shape[0]), _ptr(stims), ctypes.c_uint32(T),
                                           ctypes.c_uint32(clamp_steps), ctypes.c_uint32(use_bias), _ptr(sz),
                                           ctypes.c_uint32(len(sz)), out))
    return [e.as_dict() for e in out]


def tune_lookup(ncol: int, probes: int, T: int) -> Optional[dict]:
    """The tuning-table entry nearest to (probes, T) for ncol, None if none."""
    e = TuneEntry()
    st = lib().axon_tune_lookup(ctypes.c_uint32(ncol), ctypes.c_uint32(probes), ctypes.c_uint32(T), ctypes.byref(e))
    if st == 7:  # AXON_ERR_NOT_READY
        return None
    _check(st)
    return e.as_dict()


def tune_entries() -> list:
    L = lib()
    L.axon_tune_entries.restype = ctypes.c_uint32
    n = L.axon_tune_entries(None, ctypes.c_uint32(0))
    out = (TuneEntry * max(n, 1))()
    n = L.axon_tune_entries(out, ctypes.c_uint32(n))
    return [out[i].as_dict() for i in range(n)]


def tune_clear() -> None:
    lib().axon_tune_clear()


def tune_save(path: str) -> None:
    _check(lib().axon_tune_save(os.fsencode(path)))


def tune_load(path: str) -> None:
    _check(lib().axon_tune_load(os.fsencode(path)))


def run(args: Sequence[str], trained: Optional[Net] = None, report_path: Optional[str] = None) -> None:
    """axon_run: the reference program's whole run (main(), mm:669-1911) with
    these key=value arguments; `trained` replaces training (its epochs done,
    same training keys); the report goes to report_path (None = stdout)."""
    argv = ["axon"] + list(args)
    arr = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
    _check(lib().axon_run(len(argv), arr, trained._h if trained is not None else None,
                          os.fsencode(report_path) if report_path else None))


def mem_probe(kind: int, nbytes: int, reps: int = 5) -> float:
    """Read GB/s of a streaming pass (0), 128-byte line gathers (1) or 1 KiB tile gathers (2)."""
    g = ctypes.c_double()
    _check(lib().axon_mem_probe(ctypes.c_uint32(kind), ctypes.c_uint64(nbytes), ctypes.c_uint32(reps),
                                ctypes.byref(g)))
    return g.value


def sync_probe(kind: int, groups: int, workgroups: int, reps: int = 200) -> float:
    """Microseconds per launch of groups x workgroups storing workgroups with
    (kind 1) an arrival atomic, (2) an agent-scope release fence before it,
    (3) a full agent-scope fence, or (0) neither (axon_sync_probe)."""
    us = ctypes.c_double()
    _check(lib().axon_sync_probe(ctypes.c_uint32(kind), ctypes.c_uint32(groups), ctypes.c_uint32(workgroups),
                                 ctypes.c_uint32(reps), ctypes.byref(us)))
    return us.value


def prop_probe(net: Net, spikes: np.ndarray, reps: int = 20):
    """(microseconds per propagation launch, W3 lines one launch reads) with
    `spikes` (ncol words) as the presynaptic activity (axon_prop_probe)."""
    spikes = np.ascontiguousarray(spikes, np.uint32)
    if spikes.shape != (net.ncol,):
        raise ValueError(f"spikes must hold {net.ncol} words")
    us, lines = ctypes.c_double(), ctypes.c_uint64()
    _check(lib().axon_prop_probe(net._h, _ptr(spikes), ctypes.c_uint32(reps), ctypes.byref(us), ctypes.byref(lines)))
    return us.value, lines.value


def probe_cue(net: Net, pattern: int, nreplace: int, seed: int) -> np.ndarray:
    """Recall cue: pattern's stimulus columns with nreplace of them swapped for
    other input columns; pattern = -1 draws a random cue."""
    out = np.zeros(net.ncol, np.uint32)
    _check(lib().axon_probe_cue(net._h, ctypes.c_uint32(pattern & 0xFFFFFFFF), ctypes.c_uint32(nreplace),
                                ctypes.c_uint32(seed), _ptr(out)))
    return out


def bs_test(op: int, lo: int, hi: int, planes: Sequence[np.ndarray], mask: np.ndarray) -> list:
    """k_bstest (mm:424-434): apply inc (0), dec (1) or clip[lo,hi] (2) on the GPU."""
    p = [np.ascontiguousarray(x, np.uint32).copy() for x in planes]
    m = np.ascontiguousarray(mask, np.uint32)
    _check(lib().axon_bs_test(op, lo, hi, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), _ptr(p[3]), _ptr(m),
                              ctypes.c_size_t(m.size)))
    return p


def kwta_test(values: np.ndarray, k: int):
    """kwta_ballot_margin (mm:134-144) on rows of 32 values."""
    v = np.ascontiguousarray(values, np.uint32).reshape(-1, 32)
    n = v.shape[0]
    ball, vk, vk1 = (np.zeros(n, np.uint32) for _ in range(3))
    _check(lib().axon_kwta_test(_ptr(v), ctypes.c_uint32(n), ctypes.c_uint32(k), _ptr(ball), _ptr(vk), _ptr(vk1)))
    return ball, vk, vk1


def topk_test(drives: np.ndarray, incols: int, target: int, par: bool) -> np.ndarray:
    """k_topk_par / k_topk_serial (mm:259-352) on rows of column drives."""
    d = np.ascontiguousarray(drives, np.uint32)
    if d.ndim == 1:
        d = d[None, :]
    out = np.zeros(2 * d.shape[0], np.uint32)
    _check(lib().axon_topk_test(_ptr(d), ctypes.c_uint32(d.shape[0]), ctypes.c_uint32(d.shape[1]),
                                ctypes.c_uint32(incols), ctypes.c_uint32(target), ctypes.c_uint32(int(par)), _ptr(out)))
    return out.reshape(-1, 2)
