"""CPU checks of the drop-in boundary: libcnd.so loads, exports every entry
point include/cnd.h declares, and fails cleanly without a device."""
import ctypes as C
import os
import re

import pytest

from canard_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "cnd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cnd_[a-z_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    bound = sorted(n for n, _, _ in _lib.SIGNATURES)
    assert declared_symbols() == bound


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIBPATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of cnd_block / cnd_params have the C compiler's layout."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "cnd.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(cnd_block), sizeof(cnd_params),'
                   ' offsetof(cnd_params, cfl), offsetof(cnd_params, szco));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    want = [C.sizeof(_lib.Block), C.sizeof(_lib.Params), _lib.Params.cfl.offset, _lib.Params.szco.offset]
    assert got == want


def test_create_without_device_is_a_clean_error(monkeypatch):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from canard_amd.solver import Case, CanardError
    from canard_amd import cases
    with pytest.raises(CanardError, match="no HIP device"):
        Case([cases.periodic_block(16)], {})


def test_null_arguments_are_rejected():
    L = _lib.load()
    out = C.c_void_p()
    assert L.cnd_create(None, 1, None, 1, 0, C.byref(out)) == -1
    assert b"empty case" in L.cnd_last_error(None)
    assert L.cnd_step(None, 1) == -1
