"""pytest configuration: the ``gpu`` marker and import paths.

Tests marked ``gpu`` need an MI355X and call the HIP library through its
C-ABI; every other test runs on the CPU (oracle, fixtures, host logic, and
that the C-ABI library loads and exports its symbols).
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ehd-convection-3d_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running parity case")
    # per-test limit of pytest-timeout, above a run-wide --timeout for the long cases
    config.addinivalue_line("markers", "timeout(seconds): per-test time limit")
