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
    srcs = [os.path.join(ROOT, "oracle", f) for f in ("spk_oracle.c", "spk_quad.c", "spk_oracle.h")]
    if not os.path.exists(lib) or os.path.getmtime(lib) < max(os.path.getmtime(s) for s in srcs):
        import __graft_entry__

        __graft_entry__._build_oracle()


_ensure_oracle()


@pytest.fixture(scope="session")
def lam():
    return 633e-9
