"""Shared test set-up: import paths, the `gpu` marker, and the built libraries.

`-m gpu` tests run the HIP path on a real device and compare it with the
scalar CPU oracle (oracle/) bit for bit; `-m "not gpu"` tests pin the oracle
itself and check the host-side ABI without launching kernels.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "vx-rs_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def _ensure_built():
    """Builds the libraries only when they are absent (a fresh checkout)."""
    lib = os.path.join(ROOT, "vx-rs_amd", "vxsgm", "libvxsgm.so")
    orc = os.path.join(ROOT, "oracle", "liboracle_sgm.so")
    if not os.path.exists(orc):
        import build_oracle
        build_oracle.build()
    if not os.path.exists(lib):
        import build_lib
        build_lib.build()


_ensure_built()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def gpu_available():
    import vxsgm
    return vxsgm.device_count() > 0


@pytest.fixture(scope="session")
def ctx():
    import vxsgm
    if vxsgm.device_count() == 0:
        pytest.fail("gpu test selected but no HIP device is visible")
    c = vxsgm.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def sgm(ctx):
    import vxsgm
    s = vxsgm.SGMStereo(ctx)
    yield s
    s.close()
