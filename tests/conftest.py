import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    emu = os.environ.get("MSPLAT_EMU_LIB")
    if emu:
        # CPU emulation of the kernels under AddressSanitizer (tests/cpu_emu): the GPU tests
        # run against it on the host, to bounds-check every access before a GPU run.
        from msplat_amd import _lib
        _lib._lib = _lib.bind(emu)
