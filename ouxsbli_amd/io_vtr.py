"""On-disk formats of the reference run directory, written byte-compatibly:

* ``data/Q%05d.vtr`` -- VTK XML RectilinearGrid, appended raw Float32 blocks
  (src/print.f90:125-189, :364-382): cell-centre coordinates and rho, p and
  velocity over the whole array, ghost layers included.
* ``data/kinetic_energy.d`` (``3e12.4``: t, ke, ke/ke0) and
  ``data/entropy.d`` (``2e12.4``: t, (s0-s)/s0), interior means accumulated in
  FP32 as src/print.f90:42-122 does.
* ``recal/Q%05d.dat`` -- the restart file: Q(nx,5,ny,nz) physical conserved
  FP64, either raw stream or Fortran sequential with an ``SEQFMT01`` header
  record (3D_solver/src/main.f90:38-60, :82-93).
"""
from __future__ import annotations

import math
import pathlib
import struct

import numpy as np


def fortran_e(x: float, w: int = 12, d: int = 4) -> str:
    """Fortran ``Ew.d`` edit descriptor: 0.dddd mantissa, E+xx exponent."""
    x = float(x)
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("Infinity" if x > 0 else "-Infinity").rjust(w)
    if x == 0.0:
        return f"0.{'0' * d}E+00".rjust(w)
    sign = "-" if x < 0 else ""
    a = abs(x)
    e = math.floor(math.log10(a)) + 1
    m = a / 10.0 ** e
    mant = round(m, d)
    if mant >= 1.0:  # rounding carried into the next decade
        mant /= 10.0
        e += 1
    if mant < 0.1:
        mant *= 10.0
        e -= 1
    digits = f"{mant:.{d}f}"[1:]  # ".dddd"
    # a two-digit exponent takes the E; a three-digit one replaces it
    es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) <= 99 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    return (sign + "0" + digits + es).rjust(w)


def primitives_for_print(QJ: np.ndarray, jac: float, gamma: float):
    """make_1d_for_print3 (src/print.f90:218-243): FP64 recovery, FP32 fields.
    QJ has numpy shape (nz, ny, 5, nx)."""
    q1 = QJ[:, :, 0, :]
    rho = jac * q1
    u = QJ[:, :, 1, :] / q1
    v = QJ[:, :, 2, :] / q1
    w = QJ[:, :, 3, :] / q1
    p = (gamma - 1.0) * (jac * QJ[:, :, 4, :] - 0.5 * rho * (u ** 2 + v ** 2 + w ** 2))
    vel = np.stack([u, v, w], axis=-1)  # (nz, ny, nx, 3): component fastest
    return rho.astype(np.float32), p.astype(np.float32), vel.astype(np.float32)


def primitives_for_print2d(Q: np.ndarray, gamma: float):
    """make_1d_for_print2 (src/print.f90:192-215) from the physical state
    Q(nx,4,ny) (numpy shape (ny, 4, nx)): FP64 recovery, FP32 fields, w = 0."""
    rho = Q[:, 0, :]
    u = Q[:, 1, :] / rho
    v = Q[:, 2, :] / rho
    p = (gamma - 1.0) * (Q[:, 3, :] - 0.5 * rho * (u ** 2 + v ** 2))
    vel = np.stack([u, v, np.zeros_like(u)], axis=-1)[None]  # (1, ny, nx, 3)
    return rho[None].astype(np.float32), p[None].astype(np.float32), vel.astype(np.float32)


def write_vtr(path, x, y, z, rho, p, vel, dimension: int = 3) -> None:
    """print_xml (src/print.f90:125-189) with io = 4 (Float32)."""
    ni, nj, nk = len(x), len(y), len(z)
    io = 4
    bx, by, bz = 4 + io * ni, 4 + io * nj, 4 + io * nk
    brho = 4 + io * ni * nj * nk
    bp = brho
    bv = 4 + io * dimension * ni * nj * nk
    off1, off2 = bx, bx + by
    off3 = off2 + bz
    off4 = off3 + brho
    off5 = off4 + bp
    i4 = lambda n: f"{n:4d}"
    i12 = lambda n: f"{n:12d}"
    ext = f"0 {i4(ni - 1)} 0 {i4(nj - 1)} 0 {i4(nk - 1)}"
    head = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian">\n'
        f'  <RectilinearGrid WholeExtent="{ext}">\n'
        f'    <Piece Extent="{ext}">\n'
        '      <Coordinates>\n'
        '      <DataArray type="Float32" format="appended" offset="0"/>\n'
        f'      <DataArray type="Float32" format="appended" offset="{i12(off1)}"/>\n'
        f'      <DataArray type="Float32" format="appended" offset="{i12(off2)}"/>\n'
        '    </Coordinates>\n'
        '    <PointData>\n'
        f'      <DataArray type="Float32" format="appended" offset="{i12(off3)}" Name="rho" NumberOfComponents="1"/>\n'
        f'      <DataArray type="Float32" format="appended" offset="{i12(off4)}" Name="p" NumberOfComponents="1"/>\n'
        f'      <DataArray type="Float32" format="appended" offset="{i12(off5)}" Name="velocity" '
        f'NumberOfComponents="{dimension:1d}"/>\n'
        '      </PointData>\n'
        '    </Piece>\n'
        '  </RectilinearGrid>\n'
        '  <AppendedData encoding="raw">\n'
        '  _'
    ).encode()
    with open(path, "wb") as f:
        f.write(head)
        for nbytes, arr in ((bx, x), (by, y), (bz, z), (brho, rho), (bp, p), (bv, vel)):
            f.write(struct.pack("<i", nbytes))
            f.write(np.ascontiguousarray(arr, dtype="<f4").tobytes())
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_vtr(path):
    """Reads back a file written by :func:`write_vtr`: (x, y, z, rho, p, vel)."""
    data = pathlib.Path(path).read_bytes()
    head, _, rest = data.partition(b"  _")
    import re
    ext = [int(t) for t in re.search(rb'WholeExtent="([^"]+)"', head).group(1).split()]
    ni, nj, nk = ext[1] + 1, ext[3] + 1, ext[5] + 1
    out, pos = [], 0
    for count in (ni, nj, nk, ni * nj * nk, ni * nj * nk, 3 * ni * nj * nk):
        nb = struct.unpack_from("<i", rest, pos)[0]
        assert nb == 4 + 4 * count
        out.append(np.frombuffer(rest, dtype="<f4", count=count, offset=pos + 4))
        pos += 4 + 4 * count
    x, y, z, rho, p, vel = out
    return x, y, z, rho.reshape(nk, nj, ni), p.reshape(nk, nj, ni), vel.reshape(nk, nj, ni, 3)


def real4_mean(terms: np.ndarray, count: int) -> np.float32:
    """REAL(4) accumulation of REAL(8) terms in order, then / dble(count)
    (src/print.f90:60-68, :101-110): each addition in double precision with
    the running sum rounded to single, done by the library's host helper."""
    import ctypes
    from . import _native
    t = np.ascontiguousarray(terms, dtype=np.float64).ravel()
    acc = ctypes.c_float()
    _native.check(_native.load().ox_real4_sum(t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), t.size,
                                              ctypes.byref(acc)))
    return np.float32(np.float64(np.float32(acc.value)) / float(count))


def ke_entropy(rho, p, vel, order: int, gamma: float):
    """print_KE / print_entropy over the interior (src/print.f90:42-122) with
    the reference's expression types (io = 4): the fields are REAL(4);
    vel**2 and their sum are REAL(4), the product with 0.5d0 * rho and the
    entropy term rho * log(p * rho**(-gamma)) are REAL(8); the sums are
    REAL(4) accumulators; loops run i fastest, then j, then k."""
    h = order // 2
    sl = (slice(h, rho.shape[0] - h), slice(h, rho.shape[1] - h), slice(h, rho.shape[2] - h))
    r32 = np.ascontiguousarray(rho[sl], dtype=np.float32)
    p32 = np.ascontiguousarray(p[sl], dtype=np.float32)
    v32 = np.ascontiguousarray(vel[sl], dtype=np.float32)
    count = r32.size
    sq = (v32[..., 0] * v32[..., 0] + v32[..., 1] * v32[..., 1]) + v32[..., 2] * v32[..., 2]  # REAL(4)
    r64 = r32.astype(np.float64)
    ke = real4_mean((0.5 * r64) * sq.astype(np.float64), count)
    ent = real4_mean(r64 * np.log(p32.astype(np.float64) * (r64 ** (-gamma))), count)
    return ke, ent


class Diagnostics:
    """Appends the .d lines of one run (t = nt * step * dt)."""

    def __init__(self, data_dir: pathlib.Path):
        self.dir = pathlib.Path(data_dir)
        self.ke0 = np.float32(1.0)
        self.s0 = np.float32(1.0)

    def record(self, step: int, t: float, ke: np.float32, ent: np.float32) -> None:
        """One line of each file; t is nt * step * dt in REAL(8), stored
        REAL(4) as the reference's `real(io) t` does."""
        if step == 0:
            self.ke0, self.s0 = ke, ent
        t32 = np.float32(t)
        with open(self.dir / "kinetic_energy.d", "a") as f:
            f.write(fortran_e(t32) + fortran_e(ke) + fortran_e(np.float32(ke / self.ke0)) + "\n")
        with open(self.dir / "entropy.d", "a") as f:
            f.write(fortran_e(t32) + fortran_e(np.float32((self.s0 - ent) / self.s0)) + "\n")


def write_restart(path, Q: np.ndarray, sequential: bool = False) -> None:
    """recal/Q%05d.dat: Q(nx,5,ny,nz) FP64 physical conserved, ghosts included."""
    raw = np.ascontiguousarray(Q, dtype="<f8").tobytes()
    with open(path, "wb") as f:
        if sequential:
            for rec in (b"SEQFMT01", raw):
                f.write(struct.pack("<i", len(rec)))
                f.write(rec)
                f.write(struct.pack("<i", len(rec)))
        else:
            f.write(raw)


def read_restart(path, shape) -> np.ndarray:
    """Reads either restart layout (main.f90:38-60): SEQFMT01 sequential or raw stream."""
    data = pathlib.Path(path).read_bytes()
    n = int(np.prod(shape)) * 8
    if len(data) >= 16 and struct.unpack_from("<i", data, 0)[0] == 8 and data[4:12] == b"SEQFMT01":
        pos = 16
        rl = struct.unpack_from("<i", data, pos)[0]
        if rl != n:
            raise ValueError(f"restart record holds {rl} bytes, expected {n}")
        return np.frombuffer(data, dtype="<f8", count=n // 8, offset=pos + 4).reshape(shape).copy()
    if len(data) < n:
        raise ValueError(f"restart file holds {len(data)} bytes, expected {n}")
    return np.frombuffer(data, dtype="<f8", count=n // 8).reshape(shape).copy()
