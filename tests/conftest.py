import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def _built(path):
    return os.path.exists(os.path.join(ROOT, path))


@pytest.fixture(scope="session", autouse=True)
def _require_build():
    missing = [p for p in ("oracle/liboracle.so", "oracle/mp_oracle", "metapart_amd/libmetapart_amd.so",
                           "metapart_amd/bin/mp_synth") if not _built(p)]
    if missing:
        pytest.fail("not built (python -c 'import __graft_entry__ as g; g.build()'): %s" % ", ".join(missing))


def pytest_sessionstart(session):
    # A GPU test process holds two HIP runtimes: the one bundled with torch and
    # the system one the library links.  Some GPU tests hand torch device
    # buffers to the library, which works when torch's runtime starts first.
    if "not gpu" in (session.config.getoption("markexpr") or ""):
        return
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()
