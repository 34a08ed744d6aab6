"""The Fortran host side: the bind(C) derived types have the C layout
(CPU), and the Fortran driver program reaches the HIP library through
ISO_C_BINDING with the oracle's results (GPU)."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import compare_engforce

from mdff_amd import build as B
from mdff_amd.params import MdffEnergies, MdffParams

FC = B.FC


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
def test_fortran_types_match_c_layout(tmp_path):
    src = tmp_path / "sz.f90"
    src.write_text("""program sz
  use iso_c_binding
  use mdffgpu_mod
  type(mdffgpu_params) :: p
  type(mdffgpu_energies) :: e
  print *, c_sizeof(p), c_sizeof(e)
  print *, (loc(p%alphaES) - loc(p)), (loc(p%conv_tol_ind) - loc(p)), (loc(p%precision) - loc(p))
  print *, (loc(e%tau_coul) - loc(e)), (loc(e%rmsd) - loc(e)), (loc(e%scf_tol) - loc(e))
end program
""")
    exe = tmp_path / "sz"
    if not os.path.exists(B.OUT):
        pytest.skip("libmdffgpu.so not built")
    subprocess.run([FC, "-J", str(tmp_path), "-o", str(exe),
                    os.path.join(B.FORTRAN, "mdffgpu_mod.f90"), str(src), "-L", B.HERE,
                    "-lmdffgpu", "-Wl,-rpath," + B.HERE], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(MdffParams)
    assert vals[1] == ctypes.sizeof(MdffEnergies)
    assert vals[2] == MdffParams.alphaES.offset
    assert vals[3] == MdffParams.conv_tol_ind.offset
    assert vals[4] == MdffParams.precision.offset
    assert vals[5] == MdffEnergies.tau_coul.offset
    assert vals[6] == MdffEnergies.rmsd.offset
    assert vals[7] == MdffEnergies.scf_tol.offset


def write_driver_input(path, p, itype, A, frames):
    with open(path, "wb") as f:
        f.write(bytes(p))
        f.write(np.ascontiguousarray(itype, np.int32).tobytes())
        f.write(np.asfortranarray(A, dtype=np.float64).tobytes(order="F"))
        f.write(np.int32(len(frames)).tobytes())
        for r in frames:
            for a in range(3):
                f.write(np.ascontiguousarray(r[:, a], dtype=np.float64).tobytes())


def read_driver_output(path, n, nframes):
    data = open(path, "rb").read()
    esz = ctypes.sizeof(MdffEnergies)
    rec = (3 * n + 3 * n + 3 * n + 9 * n) * 8 + esz
    assert len(data) == rec * nframes
    out = []
    for k in range(nframes):
        b = data[k * rec:(k + 1) * rec]
        arr = np.frombuffer(b[:rec - esz], dtype=np.float64)
        f = arr[:3 * n].reshape(3, n).T
        mu = arr[3 * n:6 * n].reshape(3, n).T
        ef = arr[6 * n:9 * n].reshape(3, n).T
        efg = arr[9 * n:18 * n].reshape(3, 3, n).transpose(2, 1, 0)
        en = MdffEnergies.from_buffer_copy(b[rec - esz:])
        d = en.as_dict()
        d.update(f=f, mu=mu, ef=ef, efg=efg)
        out.append(d)
    return out


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(oracle_lib):
    import golden_cases as G
    if not os.path.exists(B.DRIVER):
        pytest.skip("Fortran driver not built")
    p, itype, A, r, meta = G.ex9_test11()
    rng = np.random.default_rng(7)
    frames = [r + 0.01 * k * rng.standard_normal(r.shape) for k in range(3)]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_driver_input(fin, p, itype, A, frames)
        subprocess.run([B.DRIVER, fin, fout], check=True, timeout=120)
        got = read_driver_output(fout, p.natm, len(frames))
    orc = oracle_lib.Oracle(p, itype, A)
    for k, g in enumerate(got):
        o = orc.engforce(frames[k], itime=k + 1)
        compare_engforce(g, o)
