"""Post-processing and output formats of a SIMPLE run (host side).

Nothing here is on the solver path: these functions read fields that a
caller has copied to the host.

* ``obstacle_forces`` / ``force_coefficients``: drag and lift on the square
  obstacle (the reference writes only CSV fields, karman_taichi.py:677-683;
  forces are this package's diagnostic).
* ``strouhal``: shedding frequency of a lift signal as f D / U.
* ``write_matrix``: one field as ``<name>.csv`` (karman_taichi.py:157-159);
  ``field_dump_names`` gives the reference's per-time-step names
  (karman_taichi.py:678-683).
* ``Display``: the reference's stacked five-field image (display.py:28-50),
  written as an 8-bit grayscale PNG.
"""
from __future__ import annotations

import struct
import zlib
from typing import Dict, Sequence, Tuple

import numpy as np


# ------------------------------------------------------------------ forces
def obstacle_forces(p, u, v, mask, dx: float, dy: float, mu: float) -> Tuple[float, float]:
    """Pressure + viscous force (fx, fy) per unit depth on the solid block of
    ``mask`` (uint8/bool, (nx+2, ny+2), 1 = solid; the block's bounding box is
    used).

    Pressure: cell-centre p of the fluid cell beside each outer face, so
    fx = sum_j (p[i0-1, j] - p[i1+1, j]) dy and fy = sum_i (p[i, j0-1] -
    p[i, j1+1]) dx.  Shear: mu * u_t / (d / 2) on each face, with u_t the
    tangential velocity interpolated to the centre of the fluid cell beside
    the face and d the cell size normal to it -- the same 2 mu / d wall model
    the momentum assembly uses on solid faces.  Top and bottom faces take u
    (drag), the upstream and downstream faces take v (lift).
    """
    m = np.asarray(mask).astype(bool)
    if not m.any():
        return 0.0, 0.0
    p = np.asarray(p, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    ii, jj = np.nonzero(m)
    i0, i1, j0, j1 = ii.min(), ii.max(), jj.min(), jj.max()
    J = slice(j0, j1 + 1)
    I = slice(i0, i1 + 1)
    fx = (p[i0 - 1, J] - p[i1 + 1, J]).sum() * dy
    fy = (p[I, j0 - 1] - p[I, j1 + 1]).sum() * dx
    ut = 0.5 * (u[i0:i1 + 1, j1 + 1] + u[i0 + 1:i1 + 2, j1 + 1])
    ub = 0.5 * (u[i0:i1 + 1, j0 - 1] + u[i0 + 1:i1 + 2, j0 - 1])
    fx += mu * (ut.sum() + ub.sum()) * dx / (0.5 * dy)
    vl = 0.5 * (v[i0 - 1, j0:j1 + 1] + v[i0 - 1, j0 + 1:j1 + 2])
    vr = 0.5 * (v[i1 + 1, j0:j1 + 1] + v[i1 + 1, j0 + 1:j1 + 2])
    fy += mu * (vl.sum() + vr.sum()) * dy / (0.5 * dx)
    return float(fx), float(fy)


def force_coefficients(fx: float, fy: float, d: float, rho: float = 1.0, u_ref: float = 1.0) -> Tuple[float, float]:
    """(Cd, Cl) = F / (rho U^2 D / 2)."""
    q = 0.5 * rho * u_ref * u_ref * d
    return fx / q, fy / q


def strouhal(sig: Sequence[float], dt: float, d: float, u_ref: float = 1.0, tail: float = 0.5) -> float:
    """f D / U of an oscillating signal (the lift), from the mean spacing of
    its upward crossings of its own mean over the last ``tail`` fraction of
    the samples (crossing times interpolated linearly).  NaN when that part
    of the signal crosses fewer than three times (less than two periods)."""
    s = np.asarray(sig, dtype=np.float64)
    s = s[int(len(s) * (1.0 - tail)):]
    if s.size < 4:
        return float("nan")
    s = s - s.mean()
    k = np.nonzero((s[:-1] < 0.0) & (s[1:] >= 0.0))[0]
    if k.size < 3:
        return float("nan")
    t = k + s[k] / (s[k] - s[k + 1])     # fractional sample index of each crossing
    period = np.diff(t).mean() * dt
    return float(d / (period * u_ref))


# ------------------------------------------------------------------ CSV fields
def write_matrix(mat, name: str) -> str:
    """``<name>.csv`` with the array's rows as lines, comma-separated
    (karman_taichi.py:157-159).  Returns the path written."""
    path = name + ".csv"
    np.savetxt(path, np.asarray(mat), delimiter=",")
    return path


def field_dump_names(time_step: int) -> Dict[str, str]:
    """Names of the per-time-step field files (karman_taichi.py:678-683):
    field -> ``timestep{n:06}_{field}``; the reference's ``udiv`` is mdiv."""
    return {f: f"timestep{time_step:06}_{f}" for f in ("pcor", "p", "u", "v", "udiv")}


# ------------------------------------------------------------------ image
def display_fields(u, v, p, pcor, mdiv) -> np.ndarray:
    """The reference's display image (display.py:28-50): cell-centred u and
    v (face averages), p, p' and mdiv on the (nx+2, ny+2) cell grid, each
    scaled to [0, 1] by (f - min) / (max - min + 1e-9), concatenated along j
    into one (nx+2, 5 (ny+2)) array indexed [x, y]."""
    p = np.asarray(p, dtype=np.float64)
    NX, NY = p.shape
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    ud = 0.5 * (u[:NX, :NY] + u[1:NX + 1, :NY])
    vd = 0.5 * (v[:NX, :NY] + v[:NX, 1:NY + 1])

    def scaled(f):
        f = np.asarray(f, dtype=np.float64)
        lo, hi = f.min(), f.max()
        return (f - lo) / (hi - lo + 1.0e-9)

    return np.concatenate([scaled(f) for f in (ud, vd, p, pcor, mdiv)], axis=1)


def image_bytes(img: np.ndarray) -> np.ndarray:
    """An [x, y] image in [0, 1] as uint8 rows of a picture (row 0 at the top,
    y upwards as on screen): value = round(255 * clip(f, 0, 1))."""
    a = np.clip(np.asarray(img, dtype=np.float64), 0.0, 1.0)
    return np.ascontiguousarray(np.rint(a * 255.0).astype(np.uint8).T[::-1])


def write_png(path: str, rows: np.ndarray) -> None:
    """8-bit grayscale PNG of a (height, width) uint8 array (no filtering)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    h, w = rows.shape

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    raw = b"".join(b"\x00" + rows[r].tobytes() for r in range(h))
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def read_png_gray(path: str) -> np.ndarray:
    """Rows of an 8-bit grayscale, unfiltered PNG written by ``write_png``."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG")
    pos, idat, w, h = 8, b"", None, None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            if depth != 8 or ctype != 0:
                raise ValueError("only 8-bit grayscale")
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, w + 1)
    if np.any(rows[:, 0] != 0):
        raise ValueError("filtered rows are not supported")
    return rows[:, 1:].copy()


class Display:
    """display.py:4-50 for a SIMPLESolver: ``display(filename)`` writes the
    stacked, min-max scaled u, v, p, p', mdiv image."""

    def __init__(self, solver, *args):
        self.solver = solver
        self.nx, self.ny = solver.nx, solver.ny

    def image(self) -> np.ndarray:
        f = self.solver.fields_numpy()
        return display_fields(f["u"], f["v"], f["p"], f["pcor"], f["mdiv"])

    def display(self, filename: str) -> None:
        write_png(filename, image_bytes(self.image()))
