"""The C-ABI library loads and exports every entry point include/sphhip.h
declares (CPU only: no compute call)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sph_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared()
    for must in ("sph_ctx_create", "sph_nnps_build", "sph_summation_density", "sph_tait",
                 "sph_momentum_av", "sph_stage_step", "sph_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from pysph_amd import _lib
    lib_path = _lib.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.skip("libsphhip.so not built")
    lib = ctypes.CDLL(lib_path)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing
    # every exported sph_ symbol is declared, and the ctypes table covers all
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True,
                         text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T sph_" in l}
    assert exported == _declared()
    assert set(_lib.EXPORTED) == _declared()


def test_binding_loads_without_device():
    from pysph_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsphhip.so not built")
    lib = _lib.load_library()
    assert b"gfx950" in lib.sph_version()
    # error paths that need no device
    assert lib.sph_ctx_create(None, 0, None) == _lib.SPH_EINVAL
    assert lib.sph_last_error(None) == b"null sph_ctx"


def test_gfx950_code_object():
    from pysph_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsphhip.so not built")
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_product_path_has_no_oracle_dependency():
    """pysph_amd must not import, load or call the oracle."""
    pkg = os.path.join(ROOT, "pysph_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text and "sphref" not in text, fn


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of sph_grid, sph_prop_list(_f32), sph_step_props,
    sph_batch_desc, sph_sources, sph_pair_params, sph_wall_frames and
    sph_nnps_dev_args have the C compiler's field offsets and sizes."""
    import ctypes
    import shutil
    import subprocess
    from pysph_amd import _lib
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    structs = {name: getattr(_lib, cls) for name, cls in
               (("sph_grid", "SphGrid"), ("sph_prop_list", "PropList"),
                ("sph_prop_list_f32", "PropListF32"), ("sph_step_props", "StepProps"),
                ("sph_batch_desc", "BatchDesc"), ("sph_sources", "Sources"),
                ("sph_pair_params", "PairParams"), ("sph_wall_frames", "WallFrames"),
                ("sph_nnps_dev_args", "NnpsDevArgs"))
               if hasattr(_lib, cls)}
    assert "sph_batch_desc" in structs
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sphhip.h"', 'int main(void){']
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call([gcc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = {}
    for ln in out:
        if ln:
            c, f, v = ln.split()
            got[(c, f)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert got[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])
