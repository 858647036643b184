import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long ensemble test")


def pytest_runtest_setup(item):
    # a GPU test runs the HIP library or fails: a missing libklft_hip.so is an
    # error, never a skip or a fallback
    if item.get_closest_marker("gpu") is not None:
        import klft_amd
        klft_amd.load()
