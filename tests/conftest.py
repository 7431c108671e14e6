import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("q-pso_amd", "oracle"):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # a fresh checkout has no built libraries: build them (CPU-only compile)
    libs = [os.path.join(REPO, "oracle", "_build", "liboracle.so"),
            os.path.join(REPO, "q-pso_amd", "lib", "libqpso.so")]
    if not all(os.path.exists(p) for p in libs):
        import subprocess
        for d in ("oracle", "q-pso_amd"):
            subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(REPO, d)])
