import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X (ROCm GPU) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running test")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


def _build_shim(name, tmp_path_factory):
    import ctypes

    src = ROOT / "tests" / "native" / f"{name}.hip"
    out = tmp_path_factory.mktemp("shim") / f"{name}.so"
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-fPIC", "-shared", f"-I{ROOT / 'include'}", str(src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        pytest.fail(f"{name} build failed:\n" + res.stderr)
    return ctypes.CDLL(str(out))


@pytest.fixture(scope="session")
def numerics_shim(tmp_path_factory):
    """Host build of the device numerics header (same source as the kernels)."""
    return _build_shim("numerics_shim", tmp_path_factory)


@pytest.fixture(scope="session")
def adjoint_shim(tmp_path_factory):
    """Host build of the residual/adjoint element operations (same source as the adjoint kernels)."""
    return _build_shim("adjoint_shim", tmp_path_factory)
