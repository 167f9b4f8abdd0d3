"""Test configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs the CPU suite (oracle vs golden vectors, host logic, the
C-ABI library loads and exports every declared symbol); `-m gpu` runs the
parity tests proper, which call the HIP library through its C ABI.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "the-rule-and-compass_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
