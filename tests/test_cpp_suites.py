"""Runs the C++ ports of the reference's backend and PDHG suites
(sovereign-solver_amd/cpptests): on the CPU over the Host backend, and with
SOV_TEST_REQUIRE_HIP=1 on the GPU, where every case also runs on the Hip
backend with both PDHG loops."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sovereign-solver_amd", "bin")


def run(name, require_hip):
    env = dict(os.environ)
    env["SOV_TEST_REQUIRE_HIP"] = "1" if require_hip else "0"
    out = subprocess.run([os.path.join(BIN, name)], capture_output=True, text=True, env=env, timeout=600)
    print(out.stdout[-4000:])
    assert out.returncode == 0, out.stdout[-4000:] + out.stderr[-2000:]


@pytest.mark.parametrize("name", ["test_backend", "test_pdhg", "test_race"])
def test_cpp_suite_host(name):
    run(name, require_hip=False)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["test_backend", "test_pdhg", "test_race"])
def test_cpp_suite_hip(gpu, name):
    run(name, require_hip=True)
