"""pytest configuration: the `gpu` marker and the repository on sys.path.

Tests marked `gpu` need an MI355X (they call libscft_hip.so's compute entry
points); everything else runs on the CPU.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and libscft_hip.so")
