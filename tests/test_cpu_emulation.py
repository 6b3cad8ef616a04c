"""Every kernel and entry point of the library, compiled against the CPU emulation of HIP in
tests/cpu_emu and run under AddressSanitizer: a small training run with relocation,
bfloat16 storage, evaluation, checkpoint, PLY export and the viewer must finish without an
out-of-bounds access, a use-after-free or a divergent barrier (the emulator aborts on
those). This bounds-checks kernels on the CPU before they run on a GPU."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpu_emu"))


def test_emulated_library_is_memory_safe(tmp_path):
    import build_emu
    exe = build_emu.build()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
    r = subprocess.run([exe, str(tmp_path), "mini"], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all scenarios passed" in r.stderr
