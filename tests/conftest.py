"""Test configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs on any host (oracle, host logic, C-ABI loading);
`-m gpu` needs an MI355X and calls the gfx950 kernels through the C ABI.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("fluid-simulation-3d_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and runs HIP kernels")
