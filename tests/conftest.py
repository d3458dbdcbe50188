import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "dg-rain_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libdgrain kernels)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (reference operators built from the reference sources)")


def _ensure_built():
    import subprocess
    lib = os.path.join(ROOT, "dg-rain_amd", "lib", "libdgrain.so")
    if not os.path.exists(lib):
        import __graft_entry__
        __graft_entry__.build_lib()
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch.device("cuda:0")


def require_ref():
    import oracle
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
