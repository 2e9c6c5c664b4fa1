import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for sub in ("rabung-atlas32_amd", "oracle", "tests"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: CPU test taking tens of seconds")
