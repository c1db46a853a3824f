"""The C-ABI library: it loads, exports every symbol axon_api.h declares, and its
host-only entry points (configuration) behave like the reference's argu() and
tunables block (src/axon.mm:660-735).  No GPU is touched here."""
import ctypes
import os
import re

import pytest

import axon_amd
import oracle
from conftest import C0, C1, CONFIG_C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "axon_api.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(axon_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = axon_amd.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version():
    assert axon_amd.lib().axon_abi_version() == 1


def test_config_struct_layout_matches_header():
    assert ctypes.sizeof(axon_amd.ConfigStruct) == (len(axon_amd.CONFIG_FIELDS) + 7) * 4
    assert ctypes.sizeof(axon_amd.PresentArgs) == 8 * 4 + 8 + 8


def test_defaults_are_the_reference_builtins_not_config_c():
    d = axon_amd.Config.defaults()
    assert (d.target, d.k, d.dthr, d.inbr, d.stimcols, d.biasup, d.contrast, d.margingate, d.epochs) == \
        (256, 6, 2, 8, 128, 0, 0, 1, 30)
    assert d.trainclamp == d.T  # contrast=0 -> trainclamp defaults to T (mm:731)


@pytest.mark.parametrize("args", [
    C0, C1, CONFIG_C, [], ["k=3", "k=9"], ["contrast=1", "clampsteps=7"], ["biasup=1"],
    ["biasup=1", "target=64", "k=24", "ncol=1024", "incols=256"], ["T=20", "trainclamp=3"], ["kk=5", "k=4x"],
])
def test_parse_matches_oracle_restatement(args):
    a = axon_amd.Config.parse(args).as_dict()
    o = oracle.parse(args)
    for f in axon_amd.CONFIG_FIELDS:
        assert a[f] == getattr(o, f), (args, f)


def test_last_occurrence_wins_and_atoi():
    c = axon_amd.Config.parse(["k=3", "k=9", "nand=5x", "targetx=1"])
    assert c.k == 9 and c.nand == 5 and c.target == 256


def test_biasup_auto_rule():
    c = axon_amd.Config.parse(["biasup=1", "target=64", "k=24"])
    pr = 64 * 24 / (4096 - 512) / 32
    assert c.biasup == int(max(2.0, (1 - pr) / pr * 1))


@pytest.mark.parametrize("bad", [["incols=4096"], ["incols=100"], ["ncol=100"], ["stimcols=600"],
                                 ["inbr=33"], ["biasshift=32"], ["stim=0"], ["maxt=0"]])
def test_validate_rejects_undefined_configs(bad):
    c = axon_amd.Config.parse(bad)
    with pytest.raises(axon_amd.AxonError) as e:
        c.validate()
    assert e.value.status == 1  # AXON_ERR_INVALID_CONFIG
    assert oracle.check(oracle.parse(bad)) != ""


def test_net_create_without_a_gpu_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(axon_amd.AxonError) as e:
        axon_amd.Net(axon_amd.Config.parse(C0))
    assert e.value.status == 4  # AXON_ERR_DEVICE, no fallback path


def test_mem_probe_rejects_bad_arguments():
    """Argument checks come before any device work (no GPU needed)."""
    import axon_amd
    for kind, nbytes, reps in ((3, 1 << 20, 1), (0, 1000, 1), (0, 1 << 20, 0), (1, 65 << 30, 1)):
        with pytest.raises(axon_amd.AxonError) as e:
            axon_amd.mem_probe(kind, nbytes, reps)
        assert e.value.status == 5


def test_sync_probe_rejects_bad_arguments():
    """Argument checks of the synchronisation probe come before any device work."""
    import axon_amd
    for kind, groups, wgs, reps in ((4, 1, 1, 1), (0, 0, 1, 1), (0, 65, 1, 1), (0, 1, 0, 1), (0, 1, 65537, 1),
                                    (0, 1, 1, 0)):
        with pytest.raises(axon_amd.AxonError) as e:
            axon_amd.sync_probe(kind, groups, wgs, reps)
        assert e.value.status == 5
