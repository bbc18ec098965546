"""The C ABI boundary (include/rtw.h) without a GPU: the library loads, exports every
declared entry point, the ctypes structs match the C layout, and argument errors are
reported through return codes."""
import ctypes as C
import os
import re
import subprocess

import pytest

from rtwca_amd import abi, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rtw.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rtw_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ("rtw_version", "rtw_context_create", "rtw_context_destroy", "rtw_last_error", "rtw_load_gaunt",
              "rtw_render", "rtw_render_dev", "rtw_xyz_to_srgb16", "rtw_render_adjoint"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = host.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing
    out = subprocess.run(["nm", "-D", "--defined-only", host.LIB_PATH], capture_output=True, text=True).stdout
    for n in _declared():
        assert re.search(r"\bT %s$" % n, out, flags=re.M), n
    assert L.rtw_version() == abi.RTW_ABI_VERSION


def test_struct_layouts_match_c(tmp_path):
    prog = tmp_path / "layout.c"
    prog.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "rtw.h"
int main(void) {
  printf("%zu %zu %zu %zu\\n", sizeof(rtw_scene), sizeof(rtw_camera), sizeof(rtw_integrator), sizeof(rtw_stats));
  printf("%zu %zu %zu %zu\\n", offsetof(rtw_scene, r_in_spin_slope), offsetof(rtw_camera, fov_vert),
         offsetof(rtw_integrator, lanes_per_ray), offsetof(rtw_stats, lanes_per_ray));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)])
    sizes, offs = [list(map(int, l.split())) for l in subprocess.check_output([str(exe)]).decode().split("\n")[:2]]
    assert sizes == [C.sizeof(abi.Scene), C.sizeof(abi.Camera), C.sizeof(abi.Integrator), C.sizeof(abi.Stats)]
    assert offs == [abi.Scene.r_in_spin_slope.offset, abi.Camera.fov_vert.offset, abi.Integrator.lanes_per_ray.offset,
                    abi.Stats.lanes_per_ray.offset]


def test_errors_without_device_are_reported():
    L = host.lib()
    h = C.c_void_p()
    rc = L.rtw_context_create(0, C.byref(h))
    if rc == 0:  # a GPU is present: argument validation instead
        L.rtw_context_destroy(h)
        pytest.skip("GPU present; covered by the gpu tests")
    assert rc in (abi.RTW_E_DEVICE, abi.RTW_E_ARG)
    assert L.rtw_last_error(None)
    assert L.rtw_render(None, None, None, 0, None, None, 0, None, None, None, None, None, None) == abi.RTW_E_ARG


def test_oracle_library_loads():
    import oracle as O
    assert O.lib() is not None
