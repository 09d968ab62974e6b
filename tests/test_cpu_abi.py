"""The C ABI surface, without a GPU: the library loads, exports every entry
point include/gfsx.h declares, the ctypes images of the structs match the C
layout, and failures come back as GFSX_FAIL with a message (the reference's
int *err convention, cbindings.cpp:9-19) instead of aborting."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gfsx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gfsx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import gridfluidsim3d_amd as g
    lib = g.load_library()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    from gridfluidsim3d_amd.gfsx import Scene, SceneGrad, StepParams, StepTape
    structs = ((Scene, "gfsx_scene"), (StepParams, "gfsx_step_params"),
               (SceneGrad, "gfsx_scene_grad"), (StepTape, "gfsx_step_tape"))
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gfsx.h"', "int main(void){"]
    for cls, cname in structs:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.split("\n") if l)
    for cls, cname in structs:
        assert int(out[cname]) == ctypes.sizeof(cls)
        for f, _ in cls._fields_:
            assert int(out["%s.%s" % (cname, f)]) == getattr(cls, f).offset, (cname, f)


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except ImportError:
        return True


@pytest.mark.skipif(not _no_gpu(), reason="checks the no-device path")
def test_no_device_fails_cleanly():
    import gridfluidsim3d_amd as g
    lib = g.load_library()
    err = ctypes.c_int(-1)
    lib.gfsx_get_device_count.restype = ctypes.c_int
    assert lib.gfsx_get_device_count(ctypes.byref(err)) == 0
    assert err.value == 0
    lib.gfsx_get_error_message.restype = ctypes.c_char_p
    assert b"device" in lib.gfsx_get_error_message()
    with pytest.raises(RuntimeError, match="gfsx_create"):
        g.Context(0)


def test_null_context_is_an_error_not_a_crash():
    import gridfluidsim3d_amd as g
    lib = g.load_library()
    err = ctypes.c_int(-1)
    lib.gfsx_synchronize.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.gfsx_synchronize(None, ctypes.byref(err))
    assert err.value == 0
    lib.gfsx_get_error_message.restype = ctypes.c_char_p
    assert b"null context" in lib.gfsx_get_error_message()
    # destroying NULL is allowed, like free(NULL)
    err.value = -1
    lib.gfsx_destroy.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.gfsx_destroy(None, ctypes.byref(err))
    assert err.value == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from gridfluidsim3d_amd import pybindings as pb
    monkeypatch.setattr(pb, "_lib", None)
    monkeypatch.setattr(pb, "library_path", lambda: str(tmp_path / "libgfsx.so"))
    with pytest.raises(pb.LibraryLoadError, match="not built"):
        pb.load_library()


def test_package_has_no_cpu_fallback():
    """The product never imports the oracle."""
    pkg = os.path.join(ROOT, "gridfluidsim3d_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.replace("oracles", ""), f


def test_dropin_toggles_report_truthfully():
    """disableOpenCL / enableOpenCL / isOpenCLEnabled of the drop-in classes
    keep and report the flag, and a call while it is cleared fails with an
    error instead of running on the device (clscalarfield.cpp:499-509,
    particleadvector.cpp:189-199); setStencil refuses unknown stencils."""
    import gridfluidsim3d_amd as g
    from oracle import oracle
    g.load_library()
    if not os.path.exists(oracle.DROPIN_SO):
        pytest.skip("drop-in checker not built (needs the reference tree at build time)")
    lib = ctypes.CDLL(oracle.DROPIN_SO)
    lib.dropin_toggles.restype = ctypes.c_int
    assert lib.dropin_toggles() == 0x3FF
