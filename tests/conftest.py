import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built native library")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# fixtures recorded by their exact optimum only; the generator rebuilds the instance
GENERATED = {"plain1000_s1": ("plain_large", 1)}


def load_golden(name):
    """A golden fixture: the reference's `.dat` file (gzip) or a generated instance."""
    from flock_amd import instances as I
    if name in GENERATED:
        fn, seed = GENERATED[name]
        return getattr(I, fn)(seed)
    return I.read_dat(os.path.join(GOLD, "dat", name + ".gz"))
