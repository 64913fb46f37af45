"""Shared pytest setup: the `gpu` marker, repo imports and GPU fixtures."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libbajo_hip.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a ROCm GPU")
    from bajo_amd._lib import LIB_PATH, lib

    assert LIB_PATH.exists(), "libbajo_hip.so is not built"
    lib()
    from bajo_amd.context import default_context

    return default_context(0)


def announce(msg: str) -> None:
    """Prints (and flushes) which GPU case is about to run."""
    print(f"[gpu-case] {msg}", flush=True)
    sys.stderr.write(f"[gpu-case] {msg}\n")
    sys.stderr.flush()
