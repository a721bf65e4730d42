"""On-disk formats against byte fixtures written by hand from the
reference's format statements (no GPU):

- the .vtr file of print_xml (src/print.f90:125-189, io = 4), opened
  unformatted/stream little-endian by print_vtk_3D (:364-382): text header with
  (i4) extents and (i12) offsets, then "  _" and six blocks, each an int32
  byte count (4 + 4 n, the reference's byte_x ... byte_v) followed by n
  Float32 values, then a line feed and the closing tags;
- Fortran E12.4 fields of the .d records (:78, :120);
- the .d means in the reference's expression types with a REAL(4) running
  sum (:60-68, :101-110);
- the restart record: raw stream, or Fortran sequential with an 8-byte
  'SEQFMT01' header record before the state record, each record framed by
  4-byte length markers (3D_solver/src/main.f90:38-60, :82-93).
"""
import struct

import numpy as np
import pytest

from ouxsbli_amd import io_vtr


def _fixture_vtr(x, y, z, rho, p, vel):
    lf = "\n"
    head = (
        '<?xml version="1.0"?>' + lf +
        '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian">' + lf +
        '  <RectilinearGrid WholeExtent="0    1 0    2 0    1">' + lf +
        '    <Piece Extent="0    1 0    2 0    1">' + lf +
        '      <Coordinates>' + lf +
        '      <DataArray type="Float32" format="appended" offset="0"/>' + lf +
        '      <DataArray type="Float32" format="appended" offset="          12"/>' + lf +
        '      <DataArray type="Float32" format="appended" offset="          28"/>' + lf +
        '    </Coordinates>' + lf +
        '    <PointData>' + lf +
        '      <DataArray type="Float32" format="appended" offset="          40" Name="rho" NumberOfComponents="1"/>'
        + lf +
        '      <DataArray type="Float32" format="appended" offset="          92" Name="p" NumberOfComponents="1"/>'
        + lf +
        '      <DataArray type="Float32" format="appended" offset="         144" Name="velocity" '
        'NumberOfComponents="3"/>' + lf +
        '      </PointData>' + lf +
        '    </Piece>' + lf +
        '  </RectilinearGrid>' + lf +
        '  <AppendedData encoding="raw">' + lf +
        '  _').encode("ascii")
    body = b""
    # byte_x = 4 + 4*2, byte_y = 4 + 4*3, byte_z = 4 + 4*2, byte_rho = byte_p = 4 + 4*12, byte_v = 4 + 4*36
    for nb, a in ((12, x), (16, y), (12, z), (52, rho), (52, p), (148, vel)):
        body += struct.pack("<i", nb) + np.asarray(a, dtype="<f4").tobytes()
    tail = (lf + '  </AppendedData>' + lf + '</VTKFile>' + lf).encode("ascii")
    return head + body + tail


def test_vtr_bytes_match_print_xml(tmp_path):
    rng = np.random.default_rng(1)
    x = np.array([0.25, 0.75], np.float32)
    y = np.array([-1.0, 0.0, 1.0], np.float32)
    z = np.array([3.0, 4.0], np.float32)
    rho = rng.random((2, 3, 2)).astype(np.float32)          # (nk, nj, ni)
    p = rng.random((2, 3, 2)).astype(np.float32)
    vel = rng.random((2, 3, 2, 3)).astype(np.float32)       # component fastest
    path = tmp_path / "Q00000.vtr"
    io_vtr.write_vtr(path, x, y, z, rho, p, vel)
    want = _fixture_vtr(x, y, z, rho.ravel(), p.ravel(), vel.ravel())
    assert path.read_bytes() == want
    # the offsets point at the int32 byte counts after "  _"
    data = path.read_bytes()
    base = data.index(b"  _") + 3
    for off, nb in ((0, 12), (12, 16), (28, 12), (40, 52), (92, 52), (144, 148)):
        assert struct.unpack_from("<i", data, base + off)[0] == nb


@pytest.mark.parametrize("value,text", [
    (0.0, "  0.0000E+00"),
    (1.0, "  0.1000E+01"),
    (0.5, "  0.5000E+00"),
    (-123.456, " -0.1235E+03"),
    (1.0e-5, "  0.1000E-04"),
    (0.99996, "  0.1000E+01"),     # rounding carries into the next decade
    (9.87654e-12, "  0.9877E-11"),
    (123456.0, "  0.1235E+06"),
    (-0.000321, " -0.3210E-03"),
    (1.5e100, "  0.1500+101"),     # three-digit exponents drop the E
])
def test_fortran_e12_4(value, text):
    assert io_vtr.fortran_e(value) == text
    assert len(io_vtr.fortran_e(value)) == 12


def test_d_records(tmp_path):
    """Uniform rho = 2, u = 1: every KE term is 0.5*2*1 = 1 exactly, so the
    mean is 1; with p = 2 the entropy term is 2 ln(2 / 2^1.4)."""
    n = 6
    rho = np.full((n, n, n), 2.0, np.float32)
    p = np.full((n, n, n), 2.0, np.float32)
    vel = np.zeros((n, n, n, 3), np.float32)
    vel[..., 0] = 1.0
    ke, ent = io_vtr.ke_entropy(rho, p, vel, order=2, gamma=1.4)
    assert ke == np.float32(1.0)
    term = np.float64(2.0) * np.log(np.float64(2.0) * np.float64(2.0) ** -1.4)
    acc = np.float32(0.0)
    for _ in range(4 ** 3):  # the 4^3 interior cells, each sum rounded to single
        acc = np.float32(np.float64(acc) + term)
    assert ent == np.float32(np.float64(acc) / 64.0)
    assert abs(float(ent) - term) < 1e-6 * abs(term)
    d = io_vtr.Diagnostics(tmp_path)
    d.record(0, 0.0, ke, ent)
    d.record(1, 2.5e-3, np.float32(0.5), ent)
    lines = (tmp_path / "kinetic_energy.d").read_text().splitlines()
    assert lines == ["  0.0000E+00  0.1000E+01  0.1000E+01", "  0.2500E-02  0.5000E+00  0.5000E+00"]
    lines = (tmp_path / "entropy.d").read_text().splitlines()
    assert lines == ["  0.0000E+00  0.0000E+00", "  0.2500E-02  0.0000E+00"]


def test_real4_running_sum():
    """A REAL(4) accumulator absorbs terms below half an ulp of the running
    sum: 1 + 4096 * 2^-25 stays 1 in single precision (the double-precision
    sum would be 1 + 2^-13)."""
    terms = np.concatenate([[1.0], np.full(4096, 2.0 ** -25)])
    assert io_vtr.real4_mean(terms, 1) == np.float32(1.0)
    # while terms of a quarter ulp accumulated first do count
    terms = np.concatenate([np.full(4096, 2.0 ** -25), [1.0]])
    assert io_vtr.real4_mean(terms, 1) == np.float32(1.0 + 2.0 ** -13)


def test_restart_records(tmp_path):
    rng = np.random.default_rng(2)
    shape = (3, 4, 5, 6)  # (nz, ny, 5, nx)
    Q = rng.standard_normal(shape)
    raw = Q.astype("<f8").tobytes()
    # stream: the raw array
    io_vtr.write_restart(tmp_path / "s.dat", Q)
    assert (tmp_path / "s.dat").read_bytes() == raw
    # sequential: [8]['SEQFMT01'][8] [n][Q][n]
    want = struct.pack("<i", 8) + b"SEQFMT01" + struct.pack("<i", 8) + \
        struct.pack("<i", len(raw)) + raw + struct.pack("<i", len(raw))
    io_vtr.write_restart(tmp_path / "q.dat", Q, sequential=True)
    assert (tmp_path / "q.dat").read_bytes() == want
    for name in ("s.dat", "q.dat"):
        np.testing.assert_array_equal(io_vtr.read_restart(tmp_path / name, shape), Q)
    # a stream file whose first bytes are not a SEQFMT01 record reads as stream
    (tmp_path / "r.dat").write_bytes(raw)
    np.testing.assert_array_equal(io_vtr.read_restart(tmp_path / "r.dat", shape), Q)
    with pytest.raises(ValueError):
        io_vtr.read_restart(tmp_path / "q.dat", (3, 4, 5, 7))
