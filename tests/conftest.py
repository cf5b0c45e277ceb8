"""Shared test setup.  Registers the `gpu` marker; builds the CPU-side
libraries on first use when they are missing (never on the GPU box, where the
prebuilt files travel with the tree)."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "project-xplus_amd"
sys.path.insert(0, str(PKG))
sys.path.insert(0, str(ROOT / "tests"))

if not (PKG / "lib" / "libxplus_amd.so").exists():
    subprocess.run(["make", "-s", "-j8", "-C", str(PKG)], check=True)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def have_gpu() -> bool:
    import xplus_amd
    return xplus_amd.device_count() > 0


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.skip("no GPU visible")
    return True
