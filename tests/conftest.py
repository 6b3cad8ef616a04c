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


def pytest_terminal_summary(terminalreporter):
    """The FP64 error of every output the GPU parity tests compared (device and FP32 oracle
    against the FP64 oracle, max |x - FP64| / max |FP64|)."""
    mod = sys.modules.get("parity_util")
    rows = mod.report_lines() if mod else []
    if rows:
        terminalreporter.section("FP64 error per output")
        for r in rows:
            terminalreporter.write_line(r)
