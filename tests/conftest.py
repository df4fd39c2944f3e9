import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def small_input(tmp_path_factory):
    """The 400-point synthetic asteroid (lmax=mmax=6, nrows=6, 7.0-7.6 h)
    on a coplanar orbit (every aspect in the ecliptic)."""
    from periodsearch_amd import synth

    p = str(tmp_path_factory.mktemp("small") / "period_search_in")
    synth.write_workload("small", p)
    return p


@pytest.fixture(scope="session")
def small_incl_input(tmp_path_factory):
    """The same asteroid on an orbit inclined by 10 deg to the ecliptic, so
    that the pole is not ambiguous between lambda and lambda+180 deg."""
    from periodsearch_amd import synth

    p = str(tmp_path_factory.mktemp("small_incl") / "period_search_in")
    synth.write_workload("small", p, incl_deg=10.0)
    return p
