"""Shared test setup.

Markers: ``gpu`` -- needs an MI355X (HIP device) and the built libsfeng.so.
The CPU oracle (oracle/) and the CPU harness (tests/harness/) are built on
first use when missing.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
HARNESS = os.path.join(ROOT, "tests", "harness")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built libsfeng.so")


def build_harness(force=False):
    so = os.path.join(HARNESS, "libcpu_harness.so")
    if force or not os.path.exists(so):
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-o", so,
                        os.path.join(HARNESS, "cpu_harness.cpp")], check=True)
    return so


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def water_sp():
    from segfault_py_amd import workloads as W
    return W.config_molecule(0)


@pytest.fixture(scope="session")
def water_ccpvdz():
    from segfault_py_amd.molecule import Molecule
    return Molecule(os.path.join(GOLDEN, "molecules", "water.xyz"),
                    os.path.join(GOLDEN, "cc-pvdz-HCNO.dat"))


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from segfault_py_amd import engine
    engine.lib()
    return torch
