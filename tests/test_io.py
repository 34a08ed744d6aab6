"""Boundary file formats (POSFF, CONTFF, TRAJFF, EFGALL, DIPFF, EFALL)
against files the reference itself wrote or reads (examples/ex0/POSFF,
examples/ex10/POSFF.start, examples/ex4/TRAJFF.*, kept under
tests/golden/files), and round trips for the rest."""
import io as _io
import os

import numpy as np
import pytest

from mdff_amd import io

FILES = os.path.join(os.path.dirname(__file__), "golden", "files")


def _f(name):
    return os.path.join(FILES, name)


def test_fortran_edit_descriptors():
    # values and their text as they appear in examples/ex10/POSFF.start
    assert io.fmt_e(-7.38624281749, 20, 12) == " -0.738624281749E+01"
    assert io.fmt_e(0.00953364968589, 20, 12) == "  0.953364968589E-02"
    assert io.fmt_e(0.0, 20, 12) == "  0.000000000000E+00"
    assert io.fmt_f(15.308620851101, 20, 12) == "     15.308620851101"
    assert io.fmt_e(2.5e-120, 20, 12) == "  0.250000000000-119"
    assert io._fnum("0.250000000000-119") == 2.5e-120
    assert io._fnum("1.5D+02") == 150.0
    assert io.fmt_e(1.0, 24, 16) == "  0.1000000000000000E+01"


def test_contff_writer_reproduces_reference_file():
    """examples/ex10/POSFF.start was written by write_CONTFF
    (src/config.f90:206-250): read it with restart_data 'rvf' and write it
    back; the text is identical byte for byte."""
    fr = io.read_posff(_f("ex10_POSFF.start"), "rvf", lrestart=True)
    assert fr.natm == 300 and fr.atypei == ["O", "B"] and fr.natmi == [180, 120]
    assert np.allclose(fr.A, 15.308620851101 * np.eye(3), rtol=0, atol=0)
    assert list(fr.itype()[[0, 179, 180, 299]]) == [1, 1, 2, 2]
    buf = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"contff_{os.getpid()}")
    try:
        io.write_contff(buf, fr)
        assert open(buf).read() == open(_f("ex10_POSFF.start")).read()
    finally:
        if os.path.exists(buf):
            os.remove(buf)


def test_posff_direct_coordinates_and_restart_flags():
    """examples/ex0/POSFF: fcc positions in Direct coordinates (dirkar, then
    wrapped into [-1/2, 1/2] cell units, src/read_pos.f90:112-120)."""
    fr = io.read_posff(_f("ex0_POSFF"))
    L = 20.993868
    assert fr.natm == 256 and fr.ntype == 1
    assert np.allclose(fr.r[1], [0.125 * L, 0.125 * L, 0.0], rtol=1e-15)
    s = fr.r / L
    assert s.min() >= -0.5 and s.max() <= 0.5
    # velocities and forces are dropped without lrestart (:99-110)
    fr2 = io.read_posff(_f("ex10_POSFF.start"), "rvf", lrestart=False)
    assert np.all(fr2.v == 0) and np.all(fr2.f == 0)


def test_trajff_reads_reference_inputs():
    fr = next(io.read_trajff(_f("ex4_TRAJFF.cluster")))
    assert fr.natm == 9 and fr.atypei == ["A", "B"] and fr.natmi == [8, 1]
    assert np.allclose(fr.r[8], [-0.62892, -0.24866, -0.73297])
    assert np.allclose(fr.A, 10.0 * np.eye(3))
    two = list(io.read_trajff(_f("ex4_TRAJFF.twocharges"), "rvf"))
    assert len(two) == 1 and np.allclose(two[0].r, [[0.5, 0, 0], [-0.5, 0, 0]])


def _frame(rng, n=7, skew=False):
    A = np.diag([9.0, 10.0, 11.0])
    if skew:
        A[0, 1], A[2, 0] = 1.5, -0.7
    s = rng.uniform(-0.5, 0.5, (n, 3))
    return io.Frame("TEST frame", A, ["Si", "O"], [3, n - 3], s @ A.T,
                    rng.normal(size=(n, 3)), rng.normal(size=(n, 3)))


@pytest.mark.parametrize("formatted", [True, False])
@pytest.mark.parametrize("data", io.DATA_KINDS)
def test_trajff_round_trip(tmp_path, formatted, data):
    rng = np.random.default_rng(3)
    frames = [_frame(rng) for _ in range(3)]
    path = str(tmp_path / "TRAJFF")
    with open(path, "w" if formatted else "wb") as fh:
        for fr in frames:
            io.write_trajff(fh, fr, data, formatted)
    back = list(io.read_trajff(path, data, formatted))
    assert len(back) == 3
    tol = 1e-11 if formatted else 0.0
    for a, b in zip(frames, back):
        assert b.natmi == a.natmi and b.atypei == a.atypei
        assert np.allclose(b.r, io.wrap(a.r, a.A), rtol=tol, atol=tol * 10)
        if data in ("rvn", "rvf"):
            assert np.allclose(b.v, a.v, rtol=tol, atol=tol)
        if data in ("rnf", "rvf"):
            assert np.allclose(b.f, a.f, rtol=tol, atol=tol)
        assert np.array_equal(b.A, a.A)


def test_trajff_cell_lines_are_rows():
    """write_trajff_xyz writes (A(i,j), j=1,3) per line while read_traj takes
    each line as a lattice vector (src/config.f90:565-567, 653-655): a
    skewed cell comes back transposed, in the reference as here."""
    rng = np.random.default_rng(5)
    fr = _frame(rng, skew=True)
    buf = _io.StringIO()
    io.write_trajff(buf, fr)
    path_text = buf.getvalue()
    tk = io._Tokens(path_text)
    A_read = io._header(tk, "TRAJFF")[2]
    assert np.allclose(A_read, fr.A.T)


@pytest.mark.parametrize("formatted", [True, False])
def test_efgall_round_trip(tmp_path, formatted):
    rng = np.random.default_rng(8)
    fr = _frame(rng)
    e = rng.normal(size=(fr.natm, 3, 3))
    e = e + e.transpose(0, 2, 1)
    path = str(tmp_path / "EFGALL")
    with open(path, "w" if formatted else "wb") as fh:
        io.write_efgall(fh, fr, e, formatted)
        io.write_efgall(fh, fr, 2 * e, formatted)
    back = list(io.read_efgall(path, formatted))
    assert len(back) == 2
    # e24.16 carries 16 significant digits; unformatted records are exact
    tol = 1e-15 if formatted else 0.0
    assert np.allclose(back[0].efg6, io.efg_six(e), rtol=tol, atol=0)
    assert np.allclose(back[1].efg6, io.efg_six(2 * e), rtol=tol, atol=0)
    assert back[0].natmi == fr.natmi and back[0].atypei == fr.atypei


def test_site_files_layout(tmp_path):
    rng = np.random.default_rng(9)
    fr = _frame(rng)
    mu = rng.normal(size=(fr.natm, 3))
    buf = _io.StringIO()
    io.write_dipff(buf, fr, mu)
    lines = buf.getvalue().splitlines()
    assert lines[8] == io.DIP_HEAD
    assert lines[9].startswith("       1  Si ") and len(lines[9]) == 8 + 2 + 3 + 3 * 24
    path = str(tmp_path / "DIPFF")
    open(path, "w").write(buf.getvalue())
    assert np.allclose(next(io.read_sites(path, 3)), mu, rtol=1e-15, atol=0)
    # EFALL skips Wannier-centre species (lwfc < 0)
    buf = _io.StringIO()
    io.write_efall(buf, fr, mu, lwfc=[-1, 0])
    assert len(buf.getvalue().splitlines()) == 9 + fr.natm - 3
