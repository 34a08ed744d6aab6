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
from mdff_amd.params import MdffBatchConfig, MdffEfgOpts, MdffEnergies, MdffParams

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
                                       (MdffBatchConfig, "mdffgpu_batch_config"),
                                       (MdffEfgOpts, "mdffgpu_efg_opts")])
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


def _create_status(p, itype=None):
    """mdffgpu_create's status and message for parameters that fail its
    checks; these run before any device work, so no GPU is needed."""
    import numpy as np
    L = _lib.lib()
    it = np.ones(max(p.natm, 1), np.int32) if itype is None else np.asarray(itype, np.int32)
    A = np.ascontiguousarray((np.eye(3) * 20.0).reshape(-1, order="F"))
    h = ctypes.c_void_p()
    st = L.mdffgpu_create(ctypes.byref(p), it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(h))
    assert not h.value
    return st, L.mdffgpu_last_error(None).decode()


def _pim_params(n=8):
    from mdff_amd import params as P
    p = P.field_defaults(n, 2)
    p.lcoulomb = 1
    p.cutlongrange = 6.0
    p.qch[0], p.qch[1] = 1.0, -1.0
    p.kES[:] = [4, 4, 4]
    p.ldip_polar[1] = 1
    P.field_check_tag(p, poldip_iso=[0.0, 1.0])
    return p


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libmdffgpu.so not built")
def test_create_failure_statuses():
    """The failure statuses of the C ABI for host-detectable errors, as the
    reference stops on them: Tang-Toennies order above 8 (src/field.f90:
    355-358) -> MDFFGPU_ERR_TT_ORDER; quadrupole tasks task_coul(4:6) in
    mixed precision, quadrupole polarizabilities with the conjugate-gradient
    SCF and polynomial extrapolation of order >= 2 ->
    MDFFGPU_ERR_UNSUPPORTED; bad sizes and non-symmetric quadrupole tables
    -> MDFFGPU_ERR_ARG."""
    from mdff_amd import params as P
    p = _pim_params()
    k = P.didx(2, 2, 1)
    p.ldip_damping[k] = 1
    p.dip_pol_damp_b[k], p.dip_pol_damp_c[k], p.dip_pol_damp_k[k] = 2.0, 1.0, 9
    st, msg = _create_status(p)
    assert st == 5 and "Tang-Toennies" in msg, (st, msg)
    for q in range(3):
        # quadrupole tasks run on the FP64 path only
        p = _pim_params()
        p.task_coul_quad[q] = 1
        p.precision = P.PREC_MIXED
        st, msg = _create_status(p)
        assert st == 6 and "quadrupole" in msg, (st, msg)
    p = _pim_params()
    p.lquad_polar[1] = 1
    P.field_check_tag(p, polquad_iso=[0.0, 0.5])
    p.scf_solver = P.SCF_PCG
    st, msg = _create_status(p)
    assert st == 6 and "Jacobi" in msg, (st, msg)
    p.scf_solver = P.SCF_JACOBI
    p.polquad[1] = 0.0  # polquad(2,1,1,1), the divisor of u_pol
    assert _create_status(p)[0] == 1
    p = _pim_params()
    p.quad[P.qidx(1, 1, 2)] = 0.3  # quad(1,1,2) without quad(1,2,1)
    st, msg = _create_status(p)
    assert st == 1 and "symmetric" in msg, (st, msg)
    p = _pim_params()
    p.algo_ext_dipole, p.extrapolate_order = P.EXT_POLY, 2
    st, msg = _create_status(p)
    assert st == 6 and "extrapolation" in msg, (st, msg)
    p = _pim_params()
    p.algo_ext_dipole, p.extrapolate_order = P.EXT_ASPC, 6
    assert _create_status(p)[0] == 1
    p = _pim_params()
    p.natm = 0
    assert _create_status(p)[0] == 1
    p = _pim_params()
    p.ntype = 17
    assert _create_status(p)[0] == 1
    p = _pim_params()
    p.kES[1] = 0
    assert _create_status(p)[0] == 1
    p = _pim_params()
    p.precision = 7
    assert _create_status(p)[0] == 1
