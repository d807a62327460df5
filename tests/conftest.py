import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "patchvvc-plus_amd"))
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding
    binding.lib()
    return binding


@pytest.fixture(scope="session")
def pvvc():
    import pvvc
    return pvvc


@pytest.fixture(scope="session")
def c0_stream(pvvc):
    """One 8-frame GoP of the 9-bit synthetic humanoid (config 0 shape)."""
    return pvvc.synth_stream("c0", n_frames=8, seed=7, threads=8)


@pytest.fixture(scope="session")
def c0_files(c0_stream, tmp_path_factory):
    d = tmp_path_factory.mktemp("c0")
    paths = []
    for i, b in enumerate(c0_stream):
        p = d / f"humanoid_{i}.frame"
        p.write_bytes(b)
        paths.append(p)
    return paths
