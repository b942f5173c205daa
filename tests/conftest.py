import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bits_equal(a, b) -> bool:
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


OPS_CASES = ["16x16x16", "24x16x16", "rough_12x10x10", "7x3x7", "1x4x4"]
TRAJ_CASES = ["16", "24x16x16", "32", "64_c1"]


def load_ops(name):
    z = np.load(GOLDEN / f"ops_{name}.npz", allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["digests"] = json.loads(str(d["digests"]))
    d["dims"] = tuple(int(x) for x in d["dims"])
    return d


def load_traj(name):
    return json.loads((GOLDEN / f"traj_{name}.json").read_text())


@pytest.hookimpl(hookwrapper=True, trylast=True)
def pytest_runtest_call(item):
    """GPU tests print their progress as they run instead of at their end, so a
    long one (the FP64 oracle of a 256^3 run takes minutes on the host) keeps
    writing to the terminal.  The wrapper runs inside the capture plugin's,
    which resumes capture for the call phase, and suspends it again.  Tests
    that read captured output keep capture."""
    if item.get_closest_marker("gpu") is None or "capfd" in getattr(item, "fixturenames", ()):
        yield
        return
    capman = item.config.pluginmanager.getplugin("capturemanager")
    with capman.global_and_fixture_disabled():
        yield


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from heterogeneous_mhd_amd import lib
    lib()  # raises if the HIP library is missing: no fallback
    return torch.device("cuda:0")
