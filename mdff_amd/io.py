"""Configuration and per-site output files of MDFF, on the host side of the
force/field boundary.

Readers and writers of the formats the reference exchanges with the
evaluation it hands to this library:

* POSFF   -- input configuration, read_pos (src/read_pos.f90:60-121)
* CONTFF  -- last configuration, write_CONTFF (src/config.f90:206-250)
* TRAJFF  -- trajectory frames, write_trajff_xyz (src/config.f90:539-628),
             read_traj_header / read_traj (src/config.f90:630-785),
             formatted or Fortran unformatted sequential
* DIPFF   -- dipoles, write_DIPFF (src/field.f90:4051-4089)
* QUADFF  -- quadrupoles, write_QUADFF (src/field.f90:4095-4136)
* EFGALL  -- EFG tensors, write_EFGALL (src/field.f90:4142-4196) and the
             standalone EFG run (src/efg.f90:502-560); read by efg_stat
             (src/efg.f90:2106-2133, 2185-2215)
* EFALL   -- electric fields, write_EFALL (src/field.f90:4202-4243)

Explicit Fortran edit descriptors (Ew.d, Fw.d, Iw, A) are reproduced byte
for byte, including the compiler's list-directed layout of the integer and
label lines that the reference's own files show (I12 per integer, labels as
consecutive A3 fields after one blank).  List-directed REAL output is
compiler-specific; those header lines are written as 17-significant-digit
E literals, which every list-directed READ accepts.

Unformatted files use the sequential record layout of the Fortran compilers
the reference builds with: each record framed by its byte length as a
little-endian int32 before and after; INTEGER is int32, REAL(dp) float64,
CHARACTER(len=n) n bytes blank-padded (system 60, labels 3, coordinate
format 1).

Positions are Cartesian (natm, 3) arrays; the cell is A with the lattice
vectors as columns (simu_cell%A).
"""
from __future__ import annotations

import dataclasses
import struct
from typing import IO, Iterator, Optional, Sequence

import numpy as np

# coord_format_allowed (src/config.f90:95, set in config_init)
COORD_FORMATS = ("Direct", "D", "Cartesian", "C")
DATA_KINDS = ("rnn", "rvn", "rnf", "rvf")


class FormatError(ValueError):
    """A file that does not follow the reference layout (the reference
    prints a message and STOPs, e.g. src/read_pos.f90:76-79)."""


@dataclasses.dataclass
class Frame:
    """One configuration as the config module holds it."""
    system: str
    A: np.ndarray                 # (3, 3), columns = lattice vectors
    atypei: list                  # species labels, ntype
    natmi: list                   # atoms per species, ntype
    r: np.ndarray                 # (natm, 3) Cartesian
    v: Optional[np.ndarray] = None
    f: Optional[np.ndarray] = None
    atype: Optional[list] = None  # per-atom labels as read (None: from blocks)

    @property
    def natm(self) -> int:
        return int(self.r.shape[0])

    @property
    def ntype(self) -> int:
        return len(self.atypei)

    def itype(self) -> np.ndarray:
        """1-based species of every atom: consecutive blocks of natmi
        atoms per species (typeinfo_init, src/read_pos.f90:185-199)."""
        return np.repeat(np.arange(1, self.ntype + 1, dtype=np.int32),
                         np.asarray(self.natmi, dtype=np.int64))

    def labels(self) -> list:
        """Per-atom labels atype(ia) = atypei(itype(ia))."""
        return [self.atypei[t - 1] for t in self.itype()]


# ---------------------------------------------------------------------------
# lattice helpers (src/lattice.f90)

def reciprocal(A: np.ndarray) -> np.ndarray:
    """B with B^T A = 1 (simu_cell%B, src/lattice.f90:60-168)."""
    return np.linalg.inv(np.asarray(A, dtype=np.float64)).T


def kardir(r: np.ndarray, A: np.ndarray) -> np.ndarray:
    """Cartesian to fractional, s_i = sum_j r_j B(j,i) (src/lattice.f90:183-213)."""
    return np.asarray(r, dtype=np.float64) @ reciprocal(A)


def dirkar(s: np.ndarray, A: np.ndarray) -> np.ndarray:
    """Fractional to Cartesian, r_j = sum_i A(j,i) s_i (src/lattice.f90:215-257)."""
    return np.asarray(s, dtype=np.float64) @ np.asarray(A, dtype=np.float64).T


def nint(x: np.ndarray) -> np.ndarray:
    """Fortran NINT: halves rounded away from zero."""
    return np.where(x >= 0.0, np.floor(x + 0.5), np.ceil(x - 0.5))


def wrap(r: np.ndarray, A: np.ndarray) -> np.ndarray:
    """kardir, periodicbc, dirkar (src/lattice.f90:269-288): fractional
    coordinates into [-1/2, 1/2]."""
    s = kardir(r, A)
    return dirkar(s - nint(s), A)


# ---------------------------------------------------------------------------
# Fortran edit descriptors

def fmt_e(x: float, w: int, d: int) -> str:
    """Ew.d: 0.ddddE+ee, d significant digits, right-justified in w."""
    x = float(x)
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
        sign = "-" if np.signbit(x) else ""
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        exp = int(e) + 1
        body = "0." + m.replace(".", "")
        if abs(exp) <= 99:
            body += f"E{'+' if exp >= 0 else '-'}{abs(exp):02d}"
        else:  # three-digit exponent: the E is dropped
            body += f"{'+' if exp >= 0 else '-'}{abs(exp):03d}"
        sign = "-" if x < 0 else ""
    s = sign + body
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def fmt_f(x: float, w: int, d: int) -> str:
    """Fw.d."""
    s = f"{float(x):.{d}f}"
    return "*" * w if len(s) > w else s.rjust(w)


def fmt_i(k: int, w: int) -> str:
    s = str(int(k))
    return "*" * w if len(s) > w else s.rjust(w)


def _ld_reals(vals) -> str:
    """A list-directed line of reals (any valid literal reads back)."""
    return " " + "  ".join(f"{float(v): .16E}" for v in vals)


def _ld_ints(vals) -> str:
    """List-directed integers as the reference's files show them (I12)."""
    return "".join(fmt_i(v, 12) for v in vals)


def _ld_labels(labels) -> str:
    """List-directed CHARACTER(len=3) items: one blank, then the padded
    labels back to back (examples/ex10/POSFF.start line 7)."""
    return " " + "".join(str(a)[:3].ljust(3) for a in labels)


# ---------------------------------------------------------------------------
# list-directed reading

class _Tokens:
    """Whitespace/comma separated values of a text file, by line or as one
    stream (list-directed READ consumes as many records as it needs)."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def line(self) -> list:
        while self.pos < len(self.lines):
            toks = self.lines[self.pos].replace(",", " ").split()
            self.pos += 1
            if toks:
                return toks
        raise FormatError("unexpected end of file")

    def raw(self) -> str:
        if self.pos >= len(self.lines):
            raise FormatError("unexpected end of file")
        s = self.lines[self.pos]
        self.pos += 1
        return s

    def take(self, count: int) -> list:
        out = []
        while len(out) < count:
            out.extend(self.line())
        # values left in the last record are skipped, as a READ statement
        # does when its list is satisfied
        return out[:count]

    def at_end(self) -> bool:
        return all(not s.strip() for s in self.lines[self.pos:])


def _fnum(tok: str) -> float:
    # Fortran exponent letters D/d/E/e, also the E-less "0.1+100" form
    t = tok.replace("D", "E").replace("d", "E")
    try:
        return float(t)
    except ValueError:
        for k in range(len(t) - 1, 0, -1):
            if t[k] in "+-" and t[k - 1] not in "Ee":
                return float(t[:k] + "E" + t[k:])
        raise FormatError(f"not a real number: {tok!r}")


def _header(tk: _Tokens, what: str):
    natm = int(tk.line()[0])
    system = tk.raw().strip()
    rows = [[_fnum(x) for x in tk.take(3)] for _ in range(3)]
    A = np.array(rows, dtype=np.float64).T  # line k = lattice vector k = column k
    ntype = int(tk.line()[0])
    atypei = tk.take(ntype) if ntype > 0 else []
    # labels written back to back ("SiO ") cannot be split; the reference
    # reads them the same way, so a mismatch is an error there too
    natmi = [int(x) for x in tk.take(ntype)]
    cpos = tk.line()[0]
    if cpos not in COORD_FORMATS:
        raise FormatError(f"{what}: line 9 should be one of {COORD_FORMATS}, found {cpos!r}")
    if sum(natmi) != natm:
        raise FormatError(f"{what}: natmi sums to {sum(natmi)}, natm is {natm}")
    return natm, system, A, list(atypei), natmi, cpos


def _ncols(data: str) -> int:
    if data not in DATA_KINDS:
        raise ValueError(f"data kind must be one of {DATA_KINDS}")
    return {"rnn": 3, "rvn": 6, "rnf": 6, "rvf": 9}[data]


def _split_cols(vals: np.ndarray, data: str):
    r = vals[:, 0:3]
    v = f = None
    if data == "rvn":
        v = vals[:, 3:6]
    elif data == "rnf":
        f = vals[:, 3:6]
    elif data == "rvf":
        v, f = vals[:, 3:6], vals[:, 6:9]
    return r, v, f


# ---------------------------------------------------------------------------
# POSFF / CONTFF

def read_posff(path: str, restart_data: str = "rnn", lrestart: bool = False) -> Frame:
    """read_pos (src/read_pos.f90:60-121).  The atom records are one
    list-directed READ over all atoms (restart_data 'rnn' 4 items, 'rvn' 7,
    'rvf' 10 per atom).  Velocities and forces are kept only with lrestart
    (:99-110); Direct positions go through dirkar (:112-116); positions end
    wrapped into the cell (:118-120)."""
    if restart_data not in ("rnn", "rvn", "rvf"):
        raise ValueError("restart_data must be 'rnn', 'rvn' or 'rvf'")
    with open(path) as fh:
        tk = _Tokens(fh.read())
    natm, system, A, atypei, natmi, cpos = _header(tk, "POSFF")
    per = 1 + {"rnn": 3, "rvn": 6, "rvf": 9}[restart_data]
    toks = tk.take(per * natm) if natm else []
    atype = [toks[k * per] for k in range(natm)]
    vals = np.array([_fnum(toks[k * per + c]) for k in range(natm) for c in range(1, per)],
                    dtype=np.float64).reshape(natm, per - 1)
    r, v, f = _split_cols(vals, "rvf" if restart_data == "rvf" else restart_data)
    if not lrestart:
        v = np.zeros_like(r)
        f = np.zeros_like(r)
    if cpos in ("Direct", "D"):
        r = dirkar(r, A)
    r = wrap(r, A)
    return Frame(system, A, atypei, natmi, r, v, f, atype)


def write_contff(path: str, fr: Frame) -> None:
    """write_CONTFF (src/config.f90:206-250): positions wrapped, velocities
    and forces (zeros when absent), one '(a,9e20.12)' record per atom."""
    r = wrap(fr.r, fr.A)
    n = fr.natm
    v = fr.v if fr.v is not None else np.zeros((n, 3))
    f = fr.f if fr.f is not None else np.zeros((n, 3))
    out = [fmt_i(n, 12), fr.system.ljust(60)[:60]]
    for k in range(3):
        out.append("".join(fmt_f(fr.A[a, k], 20, 12) for a in range(3)))
    out.append(fmt_i(fr.ntype, 4))
    out.append(_ld_labels(fr.atypei))
    out.append(_ld_ints(fr.natmi))
    out.append("Cartesian")
    labels = fr.atype if fr.atype is not None else fr.labels()
    for i in range(n):
        vals = (*r[i], *v[i], *f[i])
        out.append(str(labels[i])[:3].ljust(3) + "".join(fmt_e(x, 20, 12) for x in vals))
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


# ---------------------------------------------------------------------------
# Fortran unformatted sequential records

def _rec_write(fh: IO[bytes], payload: bytes) -> None:
    m = struct.pack("<i", len(payload))
    fh.write(m + payload + m)


def _rec_read(fh: IO[bytes]) -> Optional[bytes]:
    head = fh.read(4)
    if not head:
        return None
    if len(head) != 4:
        raise FormatError("truncated record marker")
    (n,) = struct.unpack("<i", head)
    body = fh.read(n)
    tail = fh.read(4)
    if len(body) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
        raise FormatError("corrupt unformatted record")
    return body


def _chars(s: str, n: int) -> bytes:
    return str(s)[:n].ljust(n).encode("ascii")


def _bin_header_write(fh, fr: Frame, cell_rows: bool, coord: Optional[str]) -> None:
    _rec_write(fh, struct.pack("<i", fr.natm))
    _rec_write(fh, _chars(fr.system, 60))
    for k in range(3):
        row = fr.A[k, :] if cell_rows else fr.A[:, k]
        _rec_write(fh, struct.pack("<3d", *row))
    _rec_write(fh, struct.pack("<i", fr.ntype))
    _rec_write(fh, b"".join(_chars(a, 3) for a in fr.atypei))
    _rec_write(fh, struct.pack(f"<{fr.ntype}i", *fr.natmi))
    if coord is not None:
        _rec_write(fh, _chars(coord, 1))


def _bin_header_read(fh, what: str, with_coord: bool):
    rec = _rec_read(fh)
    if rec is None:
        return None
    (natm,) = struct.unpack("<i", rec)
    system = _rec_read(fh).decode("ascii").rstrip()
    A = np.array([struct.unpack("<3d", _rec_read(fh)) for _ in range(3)]).T
    (ntype,) = struct.unpack("<i", _rec_read(fh))
    lab = _rec_read(fh).decode("ascii")
    atypei = [lab[3 * k:3 * k + 3].strip() for k in range(ntype)]
    natmi = list(struct.unpack(f"<{ntype}i", _rec_read(fh)))
    cpos = "C"
    if with_coord:
        cpos = _rec_read(fh).decode("ascii").strip()
        if cpos not in COORD_FORMATS:
            raise FormatError(f"{what}: coordinate format {cpos!r}")
    return natm, system, A, atypei, natmi, cpos


# ---------------------------------------------------------------------------
# TRAJFF

def write_trajff(fh, fr: Frame, data: str = "rnn", formatted: bool = True) -> None:
    """One frame of write_trajff_xyz (src/config.f90:539-628), positions
    wrapped.  The cell lines hold (A(i,j), j=1,3), i.e. the rows of A,
    while read_traj takes each line as a column: the reference's own round
    trip transposes a non-symmetric cell, and so does this one."""
    ncol = _ncols(data)
    r = wrap(fr.r, fr.A)
    n = fr.natm
    cols = [r]
    if data in ("rvn", "rvf"):
        cols.append(fr.v if fr.v is not None else np.zeros((n, 3)))
    if data in ("rnf", "rvf"):
        cols.append(fr.f if fr.f is not None else np.zeros((n, 3)))
    vals = np.concatenate(cols, axis=1)
    assert vals.shape[1] == ncol
    if formatted:
        out = [fmt_i(n, 6), fr.system]
        for i in range(3):
            out.append("".join(fmt_f(fr.A[i, j], 20, 12) for j in range(3)))
        out.append(fmt_i(fr.ntype, 4))
        out.append(_ld_labels(fr.atypei))
        out.append(_ld_ints(fr.natmi))
        out.append("Cartesian")
        labels = fr.atype if fr.atype is not None else fr.labels()
        for i in range(n):
            out.append(str(labels[i])[:3].ljust(3) + "".join(fmt_e(x, 20, 12) for x in vals[i]))
        fh.write("\n".join(out) + "\n")
    else:
        _bin_header_write(fh, fr, cell_rows=True, coord="C")
        _rec_write(fh, np.ascontiguousarray(vals.T, dtype="<f8").tobytes())


def read_trajff(path: str, data: str = "rnn", formatted: bool = True) -> Iterator[Frame]:
    """Frames of read_traj (src/config.f90:695-785): one READ per atom
    record (extra values on the line are ignored); Direct positions go
    through dirkar."""
    ncol = _ncols(data)
    if formatted:
        with open(path) as fh:
            tk = _Tokens(fh.read())
        while not tk.at_end():
            natm, system, A, atypei, natmi, cpos = _header(tk, "TRAJFF")
            atype, rows = [], []
            for _ in range(natm):
                toks = tk.line()
                if len(toks) < 1 + ncol:
                    raise FormatError("TRAJFF atom record too short")
                atype.append(toks[0])
                rows.append([_fnum(x) for x in toks[1:1 + ncol]])
            vals = np.array(rows, dtype=np.float64).reshape(natm, ncol)
            r, v, f = _split_cols(vals, data)
            if cpos in ("Direct", "D"):
                r = dirkar(r, A)
            yield Frame(system, A, atypei, natmi, r, v, f, atype)
    else:
        with open(path, "rb") as fh:
            while True:
                h = _bin_header_read(fh, "TRAJFF", with_coord=True)
                if h is None:
                    return
                natm, system, A, atypei, natmi, cpos = h
                vals = np.frombuffer(_rec_read(fh), dtype="<f8").reshape(ncol, natm).T.copy()
                r, v, f = _split_cols(vals, data)
                if cpos in ("Direct", "D"):
                    r = dirkar(r, A)
                yield Frame(system, A, atypei, natmi, r, v, f)


# ---------------------------------------------------------------------------
# per-site outputs

EFG_HEAD = ("      ia type                   vxx                     vyy                     vzz"
            "                     vxy                     vxz                     vyz")
EFG_HEAD_EFGCALC = ("      ia type                   vxx                   vyy                   vzz"
                    "                   vxy                   vxz                   vyz")
DIP_HEAD = "      ia type                   mux                       muy                      muz"
QUAD_HEAD = ("      ia type                 thetaxx                 thetayy                 thetazz"
             "                 thetaxy                 thetaxz                 thetayz")
EF_HEAD = "      ia type                   Ex                      Ey                      Ez"


def _text_header(fr: Frame, natm_out: int, atypei: Sequence[str], natmi: Sequence[int]) -> list:
    out = [_ld_ints([natm_out]), " " + fr.system.ljust(60)[:60]]
    for k in range(3):
        out.append(_ld_reals(fr.A[:, k]))
    out.append(_ld_ints([len(atypei)]))
    out.append(_ld_labels(atypei))
    out.append(_ld_ints(natmi))
    return out


def _site_lines(fr: Frame, cols: np.ndarray, keep: Optional[np.ndarray]) -> list:
    labels = fr.atype if fr.atype is not None else fr.labels()
    out = []
    for i in range(fr.natm):
        if keep is not None and not keep[i]:
            continue
        out.append(fmt_i(i + 1, 8) + "  " + str(labels[i])[:3].ljust(3) +
                   "".join(fmt_e(x, 24, 16) for x in cols[i]))
    return out


def efg_six(efg: np.ndarray) -> np.ndarray:
    """(natm,3,3) -> the six reported components vxx vyy vzz vxy vxz vyz."""
    return np.stack([efg[:, 0, 0], efg[:, 1, 1], efg[:, 2, 2],
                     efg[:, 0, 1], efg[:, 0, 2], efg[:, 1, 2]], axis=1)


def write_efgall(fh, fr: Frame, efg: np.ndarray, formatted: bool = True,
                 lwfc: Optional[Sequence[int]] = None, efgcalc: bool = False) -> None:
    """One EFGALL frame.  efg is (natm,3,3).  formatted: the header and
    '(i8,2x,a3,6e24.16)' records of write_EFGALL (src/field.f90:4158-4175),
    or of efgcalc with efgcalc=True (src/efg.f90:502-527: atom count and
    species without the Wannier-centre species, lwfc < 0, which must then
    be species 1).  Unformatted: header records and efg_t as one record in
    Fortran (natm,3,3) order (src/field.f90:4178-4192, src/efg.f90:530-556;
    efgcalc omits the label records when Wannier centres are present)."""
    efg = np.asarray(efg, dtype=np.float64).reshape(fr.natm, 3, 3)
    keep = None
    atypei, natmi = list(fr.atypei), list(fr.natmi)
    nwfc = 0
    if lwfc is not None:
        it = fr.itype() - 1
        keep = np.array([lwfc[t] >= 0 for t in it])
        if efgcalc:
            nwfc = int(np.sum([lwfc[t] < 0 for t in it]))
            if nwfc:
                atypei, natmi = atypei[1:], natmi[1:]
    if formatted:
        out = _text_header(fr, fr.natm - nwfc, atypei, natmi)
        out.append(EFG_HEAD_EFGCALC if efgcalc else EFG_HEAD)
        out.extend(_site_lines(fr, efg_six(efg), keep))
        fh.write("\n".join(out) + "\n")
    else:
        _rec_write(fh, struct.pack("<i", fr.natm - nwfc))
        _rec_write(fh, _chars(fr.system, 60))
        for k in range(3):
            _rec_write(fh, struct.pack("<3d", *fr.A[:, k]))
        _rec_write(fh, struct.pack("<i", len(atypei)))
        if not (efgcalc and nwfc):
            _rec_write(fh, b"".join(_chars(a, 3) for a in atypei))
            _rec_write(fh, struct.pack(f"<{len(natmi)}i", *natmi))
        _rec_write(fh, np.asfortranarray(efg).ravel(order="F").astype("<f8").tobytes())


@dataclasses.dataclass
class EfgFrame:
    natm: int
    system: str
    A: np.ndarray
    atypei: list
    natmi: list
    efg6: np.ndarray  # (natm, 6): vxx vyy vzz vxy vxz vyz


def read_efgall(path: str, formatted: bool = True) -> Iterator[EfgFrame]:
    """Frames of an EFGALL file as efg_stat reads them
    (src/efg.f90:2106-2133, 2185-2215)."""
    if formatted:
        with open(path) as fh:
            tk = _Tokens(fh.read())
        while not tk.at_end():
            natm = int(tk.line()[0])
            system = tk.raw().strip()
            A = np.array([[_fnum(x) for x in tk.take(3)] for _ in range(3)]).T
            ntype = int(tk.line()[0])
            atypei = tk.take(ntype)
            natmi = [int(x) for x in tk.take(ntype)]
            tk.raw()  # column titles
            rows = []
            for _ in range(natm):
                toks = tk.line()
                rows.append([_fnum(x) for x in toks[2:8]])
            yield EfgFrame(natm, system, A, atypei, natmi, np.array(rows).reshape(natm, 6))
    else:
        with open(path, "rb") as fh:
            while True:
                rec = _rec_read(fh)
                if rec is None:
                    return
                (natm,) = struct.unpack("<i", rec)
                system = _rec_read(fh).decode("ascii").rstrip()
                A = np.array([struct.unpack("<3d", _rec_read(fh)) for _ in range(3)]).T
                (ntype,) = struct.unpack("<i", _rec_read(fh))
                lab = _rec_read(fh).decode("ascii")
                atypei = [lab[3 * k:3 * k + 3].strip() for k in range(ntype)]
                natmi = list(struct.unpack(f"<{ntype}i", _rec_read(fh)))
                e = np.frombuffer(_rec_read(fh), dtype="<f8").reshape(3, 3, natm).transpose(2, 1, 0)
                yield EfgFrame(natm, system, A, atypei, natmi, efg_six(e))


def write_dipff(fh, fr: Frame, mu: np.ndarray) -> None:
    """write_DIPFF (src/field.f90:4051-4089): '(i8,2x,a3,6e24.16)' with the
    three dipole components."""
    out = _text_header(fr, fr.natm, fr.atypei, fr.natmi)
    out.append(DIP_HEAD)
    out.extend(_site_lines(fr, np.asarray(mu, dtype=np.float64).reshape(fr.natm, 3), None))
    fh.write("\n".join(out) + "\n")


def write_quadff(fh, fr: Frame, theta: np.ndarray) -> None:
    """write_QUADFF (src/field.f90:4095-4136)."""
    out = _text_header(fr, fr.natm, fr.atypei, fr.natmi)
    out.append(QUAD_HEAD)
    out.extend(_site_lines(fr, efg_six(np.asarray(theta).reshape(fr.natm, 3, 3)), None))
    fh.write("\n".join(out) + "\n")


def write_efall(fh, fr: Frame, ef: np.ndarray, lwfc: Optional[Sequence[int]] = None) -> None:
    """write_EFALL (src/field.f90:4202-4243): '(i8,2x,a3,3e24.16)'."""
    keep = None
    if lwfc is not None:
        keep = np.array([lwfc[t] >= 0 for t in fr.itype() - 1])
    out = _text_header(fr, fr.natm, fr.atypei, fr.natmi)
    out.append(EF_HEAD)
    out.extend(_site_lines(fr, np.asarray(ef, dtype=np.float64).reshape(fr.natm, 3), keep))
    fh.write("\n".join(out) + "\n")


def read_sites(path: str, ncol: int) -> Iterator[np.ndarray]:
    """Per-site columns of DIPFF (3), QUADFF (6), EFALL (3) or formatted
    EFGALL (6) frames, as (natm, ncol) arrays."""
    with open(path) as fh:
        tk = _Tokens(fh.read())
    while not tk.at_end():
        natm = int(tk.line()[0])
        tk.raw()
        tk.take(9)
        ntype = int(tk.line()[0])
        tk.take(ntype)
        tk.take(ntype)
        tk.raw()
        rows = [[_fnum(x) for x in tk.line()[2:2 + ncol]] for _ in range(natm)]
        yield np.array(rows).reshape(natm, ncol)
