import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "libspk_oracle.so")
    src = os.path.join(ROOT, "oracle", "spk_oracle.c")
    if not os.path.exists(lib) or os.path.getmtime(lib) < os.path.getmtime(src):
        import __graft_entry__

        __graft_entry__._build_oracle()


_ensure_oracle()


@pytest.fixture(scope="session")
def lam():
    return 633e-9
