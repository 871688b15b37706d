"""Output formats and force diagnostics (host side, no GPU).

* forces: closed forms on chosen fields (a uniform pressure gradient gives
  G * (extent between the sampled fluid cells) * D; a uniform tangential
  velocity gives the 2 mu / d wall-shear on both faces);
* Strouhal: a sampled sine of known frequency;
* write_matrix / field names: karman_taichi.py:157-159, 678-683;
* Display: display.py:28-50 layout (face averages, per-field min-max scaling,
  concatenation along j), and the PNG decoded by an independent reader (PIL).
"""
import math

import numpy as np
import pytest

from karman_taichi_amd import cases, post


def _c1_fields():
    c = cases.c1()
    m = cases.make_mask(c)
    return c, m


def test_pressure_drag_and_lift_closed_form():
    c, m = _c1_fields()
    i0, i1, j0, j1 = c.obstacle_box()
    D = c.obstacle_cells
    xc = (np.arange(c.nx + 2) - 0.5) * c.dx          # cell-centre x of row i
    yc = (np.arange(c.ny + 2) - 0.5) * c.dy
    G = 3.0
    p = 7.0 - G * xc[:, None] + 0.0 * yc[None, :]
    u = np.zeros(c.shape_u)
    v = np.zeros(c.shape_v)
    fx, fy = post.obstacle_forces(p, u, v, m, c.dx, c.dy, c.mu)
    # sampled cells i0-1 and i1+1 are (D + 1) dx apart; D rows of height dy
    assert fx == pytest.approx(G * (D + 1) * c.dx * D * c.dy, rel=1e-12)
    assert abs(fy) < 1e-12
    p = 7.0 - G * yc[None, :] + 0.0 * xc[:, None]
    fx, fy = post.obstacle_forces(p, u, v, m, c.dx, c.dy, c.mu)
    assert fy == pytest.approx(G * (D + 1) * c.dy * D * c.dx, rel=1e-12)
    assert abs(fx) < 1e-12


def test_wall_shear_closed_form():
    c, m = _c1_fields()
    D = c.obstacle_cells
    U, V = 1.5, -0.5
    p = np.zeros(c.shape_p)
    u = np.full(c.shape_u, U)
    v = np.zeros(c.shape_v)
    fx, fy = post.obstacle_forces(p, u, v, m, c.dx, c.dy, c.mu)
    # two faces of length D dx, stress mu U / (dy / 2)
    assert fx == pytest.approx(2 * D * c.dx * c.mu * U / (0.5 * c.dy), rel=1e-12)
    assert fy == 0.0
    u[:] = 0.0
    v[:] = V
    fx, fy = post.obstacle_forces(p, u, v, m, c.dx, c.dy, c.mu)
    assert fy == pytest.approx(2 * D * c.dy * c.mu * V / (0.5 * c.dx), rel=1e-12)
    assert fx == 0.0


def test_force_coefficients():
    cd, cl = post.force_coefficients(0.3, -0.1, d=0.125, rho=1.0, u_ref=1.0)
    assert cd == pytest.approx(0.3 / (0.5 * 0.125)) and cl == pytest.approx(-0.1 / (0.5 * 0.125))


def test_strouhal_of_a_sine():
    dt, d, f = 0.0025, 0.125, 1.37
    t = np.arange(6000) * dt
    sig = 0.2 + 0.8 * np.sin(2 * math.pi * f * t + 0.3)
    assert post.strouhal(sig, dt, d) == pytest.approx(f * d, rel=1e-4)
    assert math.isnan(post.strouhal(np.ones(100), dt, d))
    assert math.isnan(post.strouhal(np.sin(np.linspace(0, 3, 50)), dt, d))


def test_write_matrix_and_field_names(tmp_path):
    a = np.arange(12, dtype=np.float64).reshape(3, 4) / 7.0
    path = post.write_matrix(a, str(tmp_path / "timestep000010_p"))
    assert path.endswith("timestep000010_p.csv")
    back = np.loadtxt(path, delimiter=",")
    np.testing.assert_array_equal(back, a)
    lines = open(path).read().splitlines()
    assert len(lines) == 3 and all(len(l.split(",")) == 4 for l in lines)
    names = post.field_dump_names(20)
    assert names == {"pcor": "timestep000020_pcor", "p": "timestep000020_p", "u": "timestep000020_u",
                     "v": "timestep000020_v", "udiv": "timestep000020_udiv"}


def test_display_layout_and_png(tmp_path):
    nx, ny = 6, 4
    rng = np.random.default_rng(0)
    u = rng.standard_normal((nx + 3, ny + 2))
    v = rng.standard_normal((nx + 2, ny + 3))
    p = rng.standard_normal((nx + 2, ny + 2))
    pc = rng.standard_normal((nx + 2, ny + 2))
    md = rng.standard_normal((nx + 2, ny + 2))
    img = post.display_fields(u, v, p, pc, md)
    assert img.shape == (nx + 2, 5 * (ny + 2))
    W = ny + 2
    # block 0: u averaged to cell centres, block 1: v, then p, p', mdiv
    ud = np.array([[0.5 * (u[i, j] + u[i + 1, j]) for j in range(ny + 2)] for i in range(nx + 2)])
    vd = np.array([[0.5 * (v[i, j] + v[i, j + 1]) for j in range(ny + 2)] for i in range(nx + 2)])
    for k, f in enumerate((ud, vd, p, pc, md)):
        blk = img[:, k * W:(k + 1) * W]
        np.testing.assert_allclose(blk, (f - f.min()) / (f.max() - f.min() + 1e-9), rtol=0, atol=1e-15)
        assert blk.min() == 0.0 and 0.999999 < blk.max() < 1.0
    rows = post.image_bytes(img)
    assert rows.shape == (5 * (ny + 2), nx + 2)       # height = stacked y, width = x
    # top-left pixel: x = 0 and the largest y (the last block's top column)
    assert rows[0, 0] == int(np.rint(255 * img[0, -1]))
    assert rows[-1, -1] == int(np.rint(255 * img[-1, 0]))
    path = str(tmp_path / "timestep000001.png")
    post.write_png(path, rows)
    np.testing.assert_array_equal(post.read_png_gray(path), rows)
    from PIL import Image
    with Image.open(path) as im:
        assert im.mode == "L" and im.size == (nx + 2, 5 * (ny + 2))
        np.testing.assert_array_equal(np.asarray(im), rows)
