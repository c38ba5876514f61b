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


def test_work_records_match_work_array():
    """The numpy structured form of beep_work (work_records, for work kept as
    column arrays) has the ctypes struct's size and offsets, and gives the
    same bytes as the item-list form."""
    import numpy as np
    assert _lib.WORK_DTYPE.itemsize == C.sizeof(_lib.Work)
    for name, _t in _lib.Work._fields_:
        assert _lib.WORK_DTYPE.fields[name][1] == getattr(_lib.Work, name).offset

    class _Ctx:
        def __init__(self, h):
            self._h = C.c_void_p(h)

    rng = np.random.default_rng(1)
    n = 37
    hs = rng.integers(1, 2 ** 47, n)
    cols = [rng.integers(0, 2, n), rng.integers(0, 1000, n), rng.integers(0, 1000, n), rng.integers(0, 1000, n),
            rng.integers(0, 2 ** 32, n), rng.integers(0, 50, n), rng.integers(50, 400, n)]
    items = [(_Ctx(int(hs[i])),) + tuple(int(c[i]) for c in cols) for i in range(n)]
    rec = _lib.work_records(hs.astype(np.uint64), *cols)
    assert bytes(_lib._work_array(items)) == rec.tobytes()
    # scalars broadcast
    rec1 = _lib.work_records(hs.astype(np.uint64), 1, cols[1], 0, cols[3], cols[4], 0, cols[6])
    assert (rec1["kind"] == 1).all() and (rec1["init"] == 0).all() and (rec1["t0"] == 0).all()
