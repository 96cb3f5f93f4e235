import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# POLYQ_AMD_PKG_DIR selects another build of the package (the sanitizer run)
for p in (os.environ.get("POLYQ_AMD_PKG_DIR") or os.path.join(ROOT, "polyq-v1.0_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
