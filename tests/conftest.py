"""Shared fixtures: paths, the CPU oracle, the reference corpora."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "squares-in-squares_amd"
sys.path.insert(0, str(PKG))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the kernels through the C ABI")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.load()


@pytest.fixture(scope="session")
def corpus(tmp_path_factory):
    """The reference device tests' input corpora, regenerated with libstdc++."""
    d = tmp_path_factory.mktemp("corpus")
    exe = d / "gen_corpus"
    subprocess.run(["g++", "-O2", "-std=c++17", str(ROOT / "tests/golden/gen_corpus.cpp"), "-o", str(exe)],
                   check=True)
    subprocess.run([str(exe), str(d)], check=True)
    self_q = np.fromfile(d / "selftest.f32", dtype=np.float32).reshape(100, 50, 3)
    seeded = np.fromfile(d / "seeded.f64", dtype=np.float64).reshape(10000, 2, 3)
    return {"selftest": self_q, "seeded": seeded}


@pytest.fixture(scope="session")
def ref_certifier():
    exe = ROOT / "oracle/_ref/certifier"
    if not exe.exists():
        pytest.skip("reference certifier not built (oracle/_ref): the reference tree was absent at build time")
    return exe


@pytest.fixture(scope="session")
def sis():
    import sisamd
    sisamd.lib()
    return sisamd
