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
