"""CPU oracle binding -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle.so, the C++ restatement of the reference
engine (see oracle/oracle.h for what pins each part).  Only the tests,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as the
checker or the timed CPU baseline; the product (axon_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
CLI_PATH = os.path.join(HERE, "axon_oracle")

FIELDS = (
    "ncol incols nand target k dthr boost mode "
    "epochs T clampsteps stim seed quiet split unionchar "
    "biasup biasdown biasshift biasdownevery inbr stimcols diag profile "
    "topkpar vtopk contrast bandlo bandhi margingate hetero vbs "
    "ltdrep ltdevery ktest facil deep wave wavet trainclamp max_t"
).split()
BUFS = ("W0 W1 W2 W3 CONN SPK_A SPK_B DRIVE BBITS ACC BIAS MARG REFBALL DRIVELOG MULT "
        "STIM STATS HIST THETA").split()
U32P = ctypes.POINTER(ctypes.c_uint32)


class OraConfig(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint32) for f in FIELDS] + [("reserved", ctypes.c_uint32 * 7)]


class OraPinned(ctypes.Structure):
    _fields_ = [
        ("within_pct", ctypes.c_double), ("between_pct", ctypes.c_double), ("fin", ctypes.c_double),
        ("pin", ctypes.c_double), ("fout", ctypes.c_double), ("pout", ctypes.c_double),
        ("recall75", ctypes.c_double), ("other75", ctypes.c_double), ("margin75", ctypes.c_double),
        ("drive_t", ctypes.c_double * 40), ("drive_tt", ctypes.c_uint32),
        ("wave_ov", ctypes.c_double * 60), ("wave_tw", ctypes.c_uint32),
        ("sustained", ctypes.c_double), ("asm_jacc", ctypes.c_double), ("wstat_final", ctypes.c_uint32),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_net_create.restype = ctypes.c_void_p
        L.ora_net_create.argtypes = [ctypes.POINTER(OraConfig), ctypes.c_int]
        L.ora_net_destroy.argtypes = [ctypes.c_void_p]
        L.ora_net_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ora_present.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, U32P]
        L.ora_present.restype = ctypes.c_uint32
        L.ora_buffer.restype = U32P
        L.ora_buffer.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
        L.ora_final_spikes.restype = U32P
        L.ora_final_spikes.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.ora_stim_cols.argtypes = [ctypes.c_void_p, ctypes.c_uint32, U32P]
        L.ora_within.argtypes = [ctypes.c_void_p, U32P, U32P, U32P]
        L.ora_wstat.argtypes = [ctypes.c_void_p]
        L.ora_wstat.restype = ctypes.c_uint32
        L.ora_indeg.argtypes = [ctypes.c_void_p, U32P]
        L.ora_untouched.argtypes = [ctypes.c_void_p, U32P]
        L.ora_wdiag.argtypes = [ctypes.c_void_p, U32P, U32P, ctypes.c_uint32, ctypes.c_uint32, U32P]
        L.ora_bs_apply.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, U32P, U32P, U32P, U32P,
                                   U32P, ctypes.c_size_t]
        L.ora_topk.argtypes = [U32P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, U32P]
        L.ora_kwta.argtypes = [U32P, ctypes.c_uint32, U32P, U32P]
        L.ora_kwta.restype = ctypes.c_uint32
        L.ora_mt_nth.restype = ctypes.c_uint32
        L.ora_mt_nth.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.ora_libcxx_shuffle.argtypes = [U32P, ctypes.c_uint32, ctypes.c_uint32]
        L.ora_uniform01.restype = ctypes.c_double
        L.ora_uniform01.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.ora_work_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
        L.ora_run_protocol.argtypes = [ctypes.POINTER(OraConfig), ctypes.c_int, ctypes.POINTER(OraPinned),
                                       ctypes.c_char_p, ctypes.c_size_t]
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U32P)


def parse(args: Sequence[str]) -> OraConfig:
    argv = ["axon_oracle"] + list(args)
    arr = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
    c = OraConfig()
    lib().ora_config_parse(len(argv), arr, ctypes.byref(c))
    return c


def check(c: OraConfig) -> str:
    buf = ctypes.create_string_buffer(256)
    rc = lib().ora_config_check(ctypes.byref(c), buf, 256)
    return buf.value.decode() if rc else ""


class OracleNet:
    def __init__(self, cfg: OraConfig, threads: int = 1):
        self.cfg = cfg
        self.h = lib().ora_net_create(ctypes.byref(cfg), threads)
        if not self.h:
            raise ValueError(f"invalid oracle config: {check(cfg)}")
        self.ncol = cfg.ncol

    def close(self):
        if self.h:
            lib().ora_net_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def buf(self, name: str) -> np.ndarray:
        """Live view of an oracle buffer (copy before the next call)."""
        n = ctypes.c_size_t()
        p = lib().ora_buffer(self.h, BUFS.index(name), ctypes.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def present(self, p, keep, T, clamp_steps, learn, seed, use_bias=2, stim=None):
        s = None if stim is None else np.ascontiguousarray(stim, np.uint32)
        r = lib().ora_present(self.h, p, float(keep), T, clamp_steps, int(bool(learn)), seed, use_bias, _p(s))
        assert r == T
        return self

    def final_spikes(self, T: int) -> np.ndarray:
        return np.ctypeslib.as_array(lib().ora_final_spikes(self.h, T), shape=(self.ncol,)).copy()

    def stim_cols(self, p: int) -> np.ndarray:
        out = np.zeros(self.cfg.stimcols, np.uint32)
        lib().ora_stim_cols(self.h, p, _p(out))
        return out

    def work_counters(self, reset: bool = False):
        out = (ctypes.c_uint64 * 2)()
        lib().ora_work_counters(self.h, out, int(reset))
        return int(out[0]), int(out[1])

    def within(self, ball, post_mask=None) -> np.ndarray:
        out = np.zeros(16, np.uint32)
        b = np.ascontiguousarray(ball, np.uint32)
        pm = None if post_mask is None else np.ascontiguousarray(post_mask, np.uint32)
        lib().ora_within(self.h, _p(b), _p(pm), _p(out))
        return out

    def wstat(self) -> int:
        return lib().ora_wstat(self.h)

    def indeg(self) -> np.ndarray:
        out = np.zeros(self.ncol * 32, np.uint32)
        lib().ora_indeg(self.h, _p(out))
        return out

    def untouched(self) -> np.ndarray:
        out = np.zeros(self.ncol * 32, np.uint32)
        lib().ora_untouched(self.h, _p(out))
        return out

    def wdiag(self, memb, asm_spk, nstim, sel) -> np.ndarray:
        out = np.zeros(48, np.uint32)
        lib().ora_wdiag(self.h, _p(np.ascontiguousarray(memb, np.uint32)),
                        _p(np.ascontiguousarray(asm_spk, np.uint32)), nstim, sel, _p(out))
        return out


def bs_apply(op, lo, hi, planes, mask):
    p = [np.ascontiguousarray(x, np.uint32).copy() for x in planes]
    m = np.ascontiguousarray(mask, np.uint32)
    lib().ora_bs_apply(op, lo, hi, _p(p[0]), _p(p[1]), _p(p[2]), _p(p[3]), _p(m), m.size)
    return p


def topk(drive, incols, target, par) -> tuple:
    d = np.ascontiguousarray(drive, np.uint32)
    out = np.zeros(2, np.uint32)
    lib().ora_topk(_p(d), d.size, incols, target, int(par), _p(out))
    return int(out[0]), int(out[1])


def kwta(v, k) -> tuple:
    v = np.ascontiguousarray(v, np.uint32)
    vk = np.zeros(1, np.uint32)
    vk1 = np.zeros(1, np.uint32)
    b = lib().ora_kwta(_p(v), k, _p(vk), _p(vk1))
    return int(b), int(vk[0]), int(vk1[0])


def mt_nth(seed: int, n: int) -> int:
    return lib().ora_mt_nth(seed, n)


def libcxx_shuffle(a: np.ndarray, seed: int) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint32).copy()
    lib().ora_libcxx_shuffle(_p(a), a.size, seed)
    return a


def uniform01(seed: int, idx: int) -> float:
    return lib().ora_uniform01(seed, idx)


def run_protocol(args: Sequence[str], threads: int = 8):
    c = parse(args)
    pin = OraPinned()
    buf = ctypes.create_string_buffer(1 << 20)
    rc = lib().ora_run_protocol(ctypes.byref(c), threads, ctypes.byref(pin), buf, len(buf))
    if rc:
        raise ValueError("oracle protocol failed (invalid config?)")
    return pin, buf.value.decode()
