"""The C ABI: the library exports every function include/mdffgpu.h
declares, and the ctypes mirrors of the structs have the C layout.  No
compute call (runs without a GPU)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from mdff_amd import _lib
from mdff_amd.params import MdffBatchConfig, MdffEnergies, MdffParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mdffgpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdffgpu_\w+)\s*\(", text)))


def test_header_declares_the_prototyped_functions():
    assert set(declared_functions()) == set(_lib.PROTOTYPES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmdffgpu.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    L = ctypes.CDLL(_lib.LIB_PATH)
    for f in declared_functions():
        assert hasattr(L, f)


def _c_offsets(struct, fields):
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"%s\"\nint main(){\n" % HEADER
    src += 'printf("%%zu\\n", sizeof(%s));\n' % struct
    for f in fields:
        src += 'printf("%%zu\\n", offsetof(%s, %s));\n' % (struct, f)
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "o.c")
        exe = os.path.join(td, "o")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-o", exe, c], check=True)
        vals = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    return vals[0], vals[1:]


@pytest.mark.parametrize("py,cname", [(MdffParams, "mdffgpu_params"),
                                       (MdffEnergies, "mdffgpu_energies"),
                                       (MdffBatchConfig, "mdffgpu_batch_config")])
def test_struct_layout_matches_header(py, cname):
    names = [f[0] for f in py._fields_]
    size, offs = _c_offsets(cname, names)
    assert ctypes.sizeof(py) == size
    for n, o in zip(names, offs):
        assert getattr(py, n).offset == o, n


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.LibraryMissing):
        _lib.lib()
