"""Channel / cavity case definitions, obstacle geometry and seeded initial fields.

Host-side set-up only (no solver arithmetic lives here).  The staggered layout
is the reference's padded one (``simple.py:27-47``):

* ``u`` has shape (nx+3, ny+2); u[i, j] is the west face of cell (i, j).
  Physical faces are i = 1..nx+1 (inlet i=1, outlet i=nx+1), j = 1..ny.
* ``v`` has shape (nx+2, ny+3); v[i, j] is the south face of cell (i, j).
  Physical faces are i = 1..nx, j = 1..ny+1 (south wall j=1, north wall ny+1).
* ``p``, ``pcor``, ``mdiv`` and the obstacle mask have shape (nx+2, ny+2);
  cell (i, j), i = 1..nx, j = 1..ny, spans [(i-1)dx, i dx] x [(j-1)dy, j dy].

Arrays are i-major with j contiguous, exactly what ``ti.field.to_numpy()``
returns for the reference fields.

The channel cases follow the square-cylinder channel of the reference's
``karman_taichi.py`` (mask ``ct`` at ``karman_taichi.py:143-154``), resized to
the benchmark geometry: a 5.0 x 1.0 channel, uniform inlet u = 1, zero-gradient
outlet, no-slip walls, a square obstacle of D x D cells centred at x = 1.0 and
one cell above the centre line, Re = U D / nu with rho = 1 and U = 1.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

# Per-side kinds for the east boundary.  West/north/south always use the
# reference's Dirichlet semantics (normal velocity folded into b, tangential
# velocity as a sliding wall, simple.py:177-216).
EAST_WALL = 0      # simple.py:184-186 / :214-216 (bicgsimp.py:196 fix)
EAST_OUTFLOW = 1   # zero gradient + outlet mass-flux fix (karman_taichi.py:180-183, :586-614)


@dataclasses.dataclass
class Case:
    """One SIMPLE problem: grid, physics, relaxation and solver settings."""

    name: str
    nx: int
    ny: int
    lx: float = 5.0
    ly: float = 1.0
    rho: float = 1.0
    mu: float = 0.01
    dt: float = 0.01
    alpha_u: float = 1.0   # velocity-correction relaxation (docs/notes.md:75-77)
    alpha_p: float = 0.3   # pressure relaxation (simple.py:23)
    alpha_m: float = 0.7   # momentum blend (simple.py:25, karman_taichi.py:22)
    # bc[side] = [normal, tangential] exactly like simple.py:39
    bc: Dict[str, List[float]] = dataclasses.field(
        default_factory=lambda: {'w': [1.0, 0.0], 'e': [0.0, 0.0], 'n': [0.0, 0.0], 's': [0.0, 0.0]})
    east: int = EAST_OUTFLOW
    pin_pcor: bool = False            # bicgsimp.py:146-151 (cavity only)
    obstacle_cells: int = 0           # D in cells (0: no obstacle)
    obstacle_xc: float = 1.0          # centre x
    obstacle_yc_cells_above: float = 1.0   # centre y = ly/2 + this * dy
    re: Optional[float] = None
    n_steps: int = 1                  # time steps
    n_outer: int = 8                  # SIMPLE outer iterations per time step
    eps_mom: float = 1e-8
    eps_p: float = 1e-8
    seed: int = 0
    v_perturb: float = 1e-3
    mixed: bool = False               # FP32 Krylov vectors (FP64 scalars, true residuals and refinement)

    @property
    def dx(self) -> float:
        return self.lx / self.nx

    @property
    def dy(self) -> float:
        return self.ly / self.ny

    @property
    def shape_u(self):
        return (self.nx + 3, self.ny + 2)

    @property
    def shape_v(self):
        return (self.nx + 2, self.ny + 3)

    @property
    def shape_p(self):
        return (self.nx + 2, self.ny + 2)

    def obstacle_box(self):
        """Inclusive cell ranges (i0, i1, j0, j1) of the square obstacle, or None.

        Cell (i, j) spans [(i-1)dx, i dx]; the square's lower-left corner is
        rounded to the nearest cell edge.
        """
        if self.obstacle_cells <= 0:
            return None
        d = self.obstacle_cells
        yc = 0.5 * self.ly + self.obstacle_yc_cells_above * self.dy
        i0 = int(round(self.obstacle_xc / self.dx - 0.5 * d)) + 1
        j0 = int(round(yc / self.dy - 0.5 * d)) + 1
        return i0, i0 + d - 1, j0, j0 + d - 1

    def obstacle_diameter(self) -> float:
        """Physical obstacle height D (used for Re, drag and Strouhal)."""
        return self.obstacle_cells * self.dy


def make_mask(case: Case) -> np.ndarray:
    """uint8 (nx+2, ny+2) obstacle mask, 1 = solid cell (karman_taichi.py:143-154).

    The padding ring is NOT marked solid: domain edges are handled by index, as
    in simple.py:154-161 and :172-216.
    """
    m = np.zeros(case.shape_p, dtype=np.uint8)
    box = case.obstacle_box()
    if box is not None:
        i0, i1, j0, j1 = box
        if not (2 <= i0 and i1 <= case.nx - 2 and 2 <= j0 and j1 <= case.ny - 1):
            raise ValueError(f"obstacle {box} does not fit inside the {case.nx}x{case.ny} grid")
        m[i0:i1 + 1, j0:j1 + 1] = 1
    return m


def initial_fields(case: Case, mask: Optional[np.ndarray] = None):
    """Seeded initial state (u, v, p, u0, v0) in FP64, generated on the host.

    u = 0 except the inlet face column u[1, 1:ny+1] = bc['w'][0]; p = 0;
    v = U(-a, a) on interior v faces (i = 1..nx, j = 2..ny) from
    ``numpy.random.default_rng(seed)``, zero on faces that touch a solid cell.
    u0 = u and v0 = v.  The reference starts from zero (simple.py:257-264).
    """
    if mask is None:
        mask = make_mask(case)
    nx, ny = case.nx, case.ny
    u = np.zeros(case.shape_u)
    v = np.zeros(case.shape_v)
    p = np.zeros(case.shape_p)
    u[1, 1:ny + 1] = case.bc['w'][0]
    if case.v_perturb > 0.0:
        rng = np.random.default_rng(case.seed)
        pert = rng.uniform(-case.v_perturb, case.v_perturb, size=(nx, ny - 1))
        v[1:nx + 1, 2:ny + 1] = pert
        solid_face = (mask[1:nx + 1, 1:ny] | mask[1:nx + 1, 2:ny + 1]).astype(bool)
        v[1:nx + 1, 2:ny + 1][solid_face] = 0.0
    return u, v, p, u.copy(), v.copy()


def channel(name: str, nx: int, ny: int, d_cells: int, re: float, dt: float,
            n_steps: int, **kw) -> Case:
    """Square-cylinder channel: nu = U D / Re with U = 1, rho = 1."""
    c = Case(name=name, nx=nx, ny=ny, obstacle_cells=d_cells, dt=dt, re=re,
             n_steps=n_steps, **kw)
    c.mu = c.rho * 1.0 * c.obstacle_diameter() / re
    return c


def c1(**kw) -> Case:
    """320x64 channel, 8x8 cylinder, Re=150, dt=0.01, 200 steps (CPU reference case)."""
    return channel("c1_320x64_re150", 320, 64, 8, 150.0, 0.01, 200, **kw)


def c2(**kw) -> Case:
    """1280x256 channel, 32x32 cylinder, Re=150, dt=0.0025, 500 steps (benchmark case)."""
    return channel("c2_1280x256_re150", 1280, 256, 32, 150.0, 0.0025, 500, **kw)


def c5(**kw) -> Case:
    """4096x820 channel, 96x96 cylinder, Re=200, dt=0.001, 300 steps (large case)."""
    return channel("c5_4096x820_re200", 4096, 820, 96, 200.0, 0.001, 300, **kw)


def batched_cases(n: int = 96, seed: int = 0) -> List[Case]:
    """Uneven batch: nx = 256 + 64 k (k ~ U{0..28}), ny = round(nx/5),
    D = round(ny/8), Re ~ U[80, 300], dt = 0.64 dx (CFL 0.64), seeded; about
    half the cases run with FP32 Krylov vectors (``mixed``, drawn from a
    second generator seeded seed + 1 so the geometry draws stay as they are)."""
    rng = np.random.default_rng(seed)
    prec = np.random.default_rng(seed + 1)
    out = []
    for b in range(n):
        nx = 256 + 64 * int(rng.integers(0, 29))
        ny = int(round(nx / 5))
        d = max(2, int(round(ny / 8)))
        re = float(rng.uniform(80.0, 300.0))
        dt = 0.64 * (5.0 / nx)
        out.append(channel(f"batch{b:02d}_{nx}x{ny}", nx, ny, d, re, dt, 200, seed=seed + b,
                           mixed=bool(prec.random() < 0.5)))
    return out


def cavity(n: int = 32, re: float = 100.0, **kw) -> Case:
    """Lid-driven cavity as in simple.py:287-303 / bicgsimp.py:355-371:
    unit square, lid u = 1 on the north wall, mu = 1/Re, p' pinned at (1, 1)."""
    c = Case(name=f"cavity{n}_re{re:g}", nx=n, ny=n, lx=1.0, ly=1.0, mu=1.0 / re,
             dt=1e12, east=EAST_WALL, pin_pcor=True,
             bc={'w': [0.0, 0.0], 'e': [0.0, 0.0], 'n': [0.0, 1.0], 's': [0.0, 0.0]},
             alpha_u=0.8, alpha_p=0.1, alpha_m=0.5, v_perturb=0.0, **kw)
    return c


CASES = {"c1": c1, "c2": c2, "c5": c5}


def _selfcheck():  # pragma: no cover - quick manual check
    c = c1()
    print(c.obstacle_box(), c.mu, math.isclose(c.mu, 0.125 / 150))


if __name__ == "__main__":  # pragma: no cover
    _selfcheck()
