import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "sovereign-solver_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


def _libs_built():
    return all(os.path.exists(os.path.join(ROOT, p)) for p in (
        "sovereign-solver_amd/lib/libsov_hip.so", "sovereign-solver_amd/lib/libsovereign.so", "oracle/liboracle.so"))


def pytest_collection_modifyitems(config, items):
    if not _libs_built():
        raise pytest.UsageError("native libraries are not built: run python -c 'import __graft_entry__ as g; g.build()'")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_pdhg.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu():
    """The HIP C-ABI library with a visible device; fails (not skips) without one."""
    import sovereign_amd as S
    n = S.hip_device_count()
    if n < 1:
        pytest.fail("a gpu-marked test ran without a visible HIP device")
    return S.hip()
