"""The C-ABI library loads without a GPU and exports every function include/*.h declares;
the struct layouts the ctypes binding uses match the headers (sizes and offsets compiled
by the host compiler). No compute calls: those need the GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

from msplat_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
HEADERS = sorted(f for f in os.listdir(INC) if f.endswith(".h"))


def _declared_functions(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"static\s+inline[^{]*\{.*?\n\}", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        if name.startswith("msplat_"):
            names.add(name)
    return names


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    missing = []
    total = 0
    for h in HEADERS:
        for name in _declared_functions(open(os.path.join(INC, h)).read()):
            total += 1
            if not hasattr(L, name):
                missing.append(f"{h}:{name}")
    assert total >= 70, total
    assert not missing, missing


def test_reference_c_api_surface_is_complete():
    """All 34 functions of the reference's msplat_c_api.h (core/include/msplat_c_api.h)."""
    names = _declared_functions(open(os.path.join(INC, "msplat_c_api.h")).read())
    assert len(names) == 34, sorted(names)
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n


def test_struct_layouts_match_headers(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r"""
#include <stddef.h>
#include <stdio.h>
#include "msplat_c_api.h"
#include "msplat_types.h"
#include "msplat_kernels.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu\n", sizeof(MsplatConfig), offsetof(MsplatConfig, keepCrs),
           offsetof(MsplatConfig, downscaleFactor), offsetof(MsplatConfig, bgColor), sizeof(MsplatStats));
    printf("%zu %zu %zu\n", sizeof(MsplatEvalMetrics), sizeof(MsplatPixelBuffer), sizeof(MsplatCamera));
    printf("%zu %zu %zu\n", sizeof(MsplatStepParams), offsetof(MsplatStepParams, sgld_seed), sizeof(MsplatDeviceState));
    MsplatConfig c = msplat_default_config();
    printf("%d %d %d %.4f %d %.1f %.3f %.3f %.2f %d %.2f %.4f %.4f %.4f\n", c.iterations, c.shDegree,
           c.shDegreeInterval, c.ssimWeight, c.capMax, c.noiseLr, c.opacityReg, c.scaleReg, c.cullRadius,
           (int)c.keepCrs, c.downscaleFactor, c.bgColor[0], c.bgColor[1], c.bgColor[2]);
    return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", INC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    a = [int(x) for x in out[0].split()]
    assert a == [C.sizeof(_lib.MsplatConfig), _lib.MsplatConfig.keepCrs.offset,
                 _lib.MsplatConfig.downscaleFactor.offset, _lib.MsplatConfig.bgColor.offset,
                 C.sizeof(_lib.MsplatStats)]
    # the reference's layout: 9 four-byte fields, a bool at 36, float at 40, float[3] at 44
    assert a[:4] == [56, 36, 40, 44]
    b = [int(x) for x in out[1].split()]
    assert b == [C.sizeof(_lib.MsplatEvalMetrics), C.sizeof(_lib.MsplatPixelBuffer), C.sizeof(_lib.MsplatCamera)]
    c = [int(x) for x in out[2].split()]
    assert c == [C.sizeof(_lib.MsplatStepParams), _lib.MsplatStepParams.sgld_seed.offset,
                 C.sizeof(_lib.MsplatDeviceState)]
    d = _lib.default_config()
    vals = out[3].split()
    assert vals == ["30000", "3", "1000", "0.2000", "1000000", "500000.0", "0.010", "0.010", "3.00", "0", "1.00",
                    "0.6130", "0.0101", "0.3984"]
    assert (d.iterations, d.shDegree, d.capMax) == (30000, 3, 1000000)


def test_dataset_errors_do_not_cross_the_boundary():
    """Bad paths return NULL instead of throwing through C (no GPU needed)."""
    L = _lib.lib()
    assert not L.msplat_dataset_create(b"synthetic:nope", 1.0, False, 8)
    assert not L.msplat_dataset_create(b"/nonexistent/colmap", 1.0, False, 8)
    assert not L.msplat_viewer_create(b"/nonexistent.ply", 64, 64, 0.8)
    assert not L.msplat_trainer_create(None, _lib.default_config())
    assert L.msplat_dataset_num_train(None) == 0
    assert L.msplat_trainer_iteration(None) == 0


def test_bucket_wave_choice_fits_lds():
    """The level-1 bucket kernel's launch never asks for more LDS than the device allows: the
    choice counts its whole dynamic LDS (each wave, one chunk, keeps 4 B of cursor per
    block). Four waves by default wherever they fit. Band edges of the LDS formula at 160 KiB
    and 64 KiB limits."""
    L = _lib.lib()

    def lds(T, w):
        return ((T + 3) & ~3) * 4 * w
    for limit in (65536, 163840):
        for T in list(range(1, 60)) + list(range(4090, 4100)) + list(range(8100, 8200)) + list(range(10200, 10300)) + list(range(16350, 16400)) + [40000, 40960, 41000]:
            for req in (0, 1, 2, 4):
                w = L.msplat_list_bucket_waves_for(T, limit, req)
                if w == 0:
                    assert lds(T, 1) > limit
                    continue
                assert lds(T, w) <= limit, (T, limit, req, w)
                if req == 0:  # four waves wherever they fit
                    assert w == (4 if lds(T, 4) <= limit else 2 if lds(T, 2) <= limit else 1)
                if req == 2:
                    assert w == (2 if lds(T, 2) <= limit else 1)
                if req == 1:
                    assert w == 1
    assert L.msplat_list_bucket_waves_for(10240, 163840, 0) == 4
    assert L.msplat_list_bucket_waves_for(10241, 163840, 0) == 2
    assert L.msplat_list_bucket_waves_for(40960, 163840, 0) == 1
    assert L.msplat_list_bucket_waves_for(40961, 163840, 0) == 0
    assert L.msplat_list_bucket_waves_for(8160, 65536, 0) == 2
    assert L.msplat_list_bucket_waves_for(425, 163840, 0) == 4
