import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-scale CPU oracle run (tens of seconds)")


# Configuration C (reference README.md:45-47) and the small CPU-runnable case:
# ncol=512 incols=128 stimcols=64, stim=12, epochs=10, seed 0.
CONFIG_C = ("nand=4 target=64 k=24 dthr=1 mode=1 inbr=4 stimcols=64 contrast=1 margingate=0 "
            "ltdrep=1 ltdevery=2 biasup=73 biasdown=1 T=10 clampsteps=5").split()
C0 = CONFIG_C + "ncol=512 incols=128 stim=12 epochs=10 seed=0".split()
C1 = CONFIG_C + "stim=12 epochs=40 deep=1 wave=1 wavet=35 seed=0".split()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle
