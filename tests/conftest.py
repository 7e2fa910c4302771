import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURE_NAMES = ["pyramid12", "pyramid40", "mixed", "overflow", "bounce", "ccd", "mixed_ccd", "events", "sleep",
                 "features", "asleep"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


_cache = {}


def load_fixture(name):
    """(spec bytes, [Snapshot...], raw snapshot bytes per step) of a golden fixture."""
    if name in _cache:
        return _cache[name]
    from forge2d_amd import snapshot as S

    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    spec = z["spec"].tobytes()
    raw = z["snaps"].tobytes()
    snaps = S.parse_all(raw)
    raws = []
    off = 0
    for s in snaps:
        raws.append(raw[off:off + s.nbytes])
        off += s.nbytes
    _cache[name] = (spec, snaps, raws)
    return _cache[name]


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no HIP device")
    return True
