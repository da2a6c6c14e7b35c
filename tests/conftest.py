import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def workdir(tmp_path_factory):
    return tmp_path_factory.mktemp("phx")


@pytest.fixture(scope="session")
def central_workload(workdir):
    from parahpop_amd import synth
    cfg, tg = synth.make_workload(str(workdir / "w1"), 1)
    return cfg, synth.load_targets(tg)


@pytest.fixture(scope="session")
def full_workload(workdir):
    """Full force model, 70x70 field, 128 objects (seed 2 population), 1 h."""
    from parahpop_amd import synth
    cfg, tg = synth.make_workload(str(workdir / "w2"), 2, n=128, duration=3600.0)
    return cfg, synth.load_targets(tg)


@pytest.fixture(scope="session")
def gpu_lib():
    import parahpop_amd
    return parahpop_amd.library()
