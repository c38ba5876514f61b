"""The ctypes mirror of the C ABI (beep_amd/_lib.py) against include/beep_gpu.h:
a small C program compiled here with gcc prints the size of every struct and
the offset of every field, which must equal the ctypes layout.  A field added
to a header struct without the binding (or in another order) fails here,
before any call passes a misread struct."""
import ctypes as C
import json
import os
import shutil
import subprocess

import pytest

from beep_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# ctypes class -> (C type, {ctypes field: C field})
STRUCTS = {
    _lib.Prior: ("beep_prior", {}),
    _lib.Comp: ("beep_comp", {}),
    _lib.Trans: ("beep_trans", {"from_": "from"}),
    _lib.Spline: ("beep_spline", {}),
    _lib.Dirichlet: ("beep_dirichlet", {}),
    _lib.Covar: ("beep_covar", {}),
    _lib.Obs: ("beep_obs", {}),
    _lib.AgeMod: ("beep_agemod", {}),
    _lib.Mod: ("beep_mod", {}),
    _lib.Spec: ("beep_spec", {}),
    _lib.Work: ("beep_work", {}),
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    d = tmp_path_factory.mktemp("abi")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "beep_gpu.h"', "int main(void) {", 'printf("{");']
    first = True
    for cls, (ctype, ren) in STRUCTS.items():
        fields = [ren.get(f, f) for f, _ in cls._fields_]
        parts = [f'\\"size\\": %zu'] + [f'\\"{f}\\": %zu' for f in fields]
        args = [f"sizeof({ctype})"] + [f"offsetof({ctype}, {f})" for f in fields]
        sep = "" if first else ","
        first = False
        lines.append(f'printf("{sep}\\"{ctype}\\": {{{", ".join(parts)}}}", {", ".join(args)});')
    lines += ['printf("}\\n");', "return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines))
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return json.loads(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)


@pytest.mark.parametrize("cls", list(STRUCTS), ids=[v[0] for v in STRUCTS.values()])
def test_struct_layout_matches_header(c_layout, cls):
    ctype, ren = STRUCTS[cls]
    lay = c_layout[ctype]
    assert C.sizeof(cls) == lay["size"], ctype
    for f, _ in cls._fields_:
        assert getattr(cls, f).offset == lay[ren.get(f, f)], f"{ctype}.{f}"
    # Every field of the C struct is bound: the ctypes fields cover the size.
    last = cls._fields_[-1][0]
    assert getattr(cls, last).offset + getattr(cls, last).size <= lay["size"] < \
        getattr(cls, last).offset + getattr(cls, last).size + 8
