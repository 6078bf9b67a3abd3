"""The C++20 host adapter (include/cfdhip.hpp) through its test driver."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "test_host_adapter")


def run(mode):
    assert os.path.exists(DRIVER), "build the driver with cfdapp_amd/build.sh"
    p = subprocess.run([DRIVER, mode], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ok" in p.stdout


def test_factory_contract_and_validation():
    run("nodevice")


@pytest.mark.gpu
def test_adapter_on_device():
    run("device")
