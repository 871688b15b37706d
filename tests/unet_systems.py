"""Small symmetric 5-point systems with identity rows for the network tests."""
import numpy as np


def small_system(seed=0, nx=14, ny=10):
    """A symmetric 5-point system with identity rows (an obstacle and an outlet column)."""
    rng = np.random.default_rng(seed)
    NX, NY = nx + 2, ny + 2
    kx = np.zeros((NX + 1, NY)); ky = np.zeros((NX, NY + 1))
    kx[2:NX - 1, 1:NY - 1] = rng.random((NX - 3, NY - 2)) + 0.5    # face between i-1 and i
    ky[1:NX - 1, 2:NY - 1] = rng.random((NX - 2, NY - 3)) + 0.5    # face between j-1 and j
    ident = np.zeros((NX, NY), bool)
    ident[5:7, 4:6] = True
    ident[nx, 1:ny + 1] = True
    coef = np.zeros((5, NX, NY))
    for i in range(1, NX - 1):
        for j in range(1, NY - 1):
            if ident[i, j]:
                coef[0, i, j] = 1.0
                continue
            links = [(1, i - 1, j, kx[i, j]), (2, i + 1, j, kx[i + 1, j]),
                     (3, i, j + 1, ky[i, j + 1]), (4, i, j - 1, ky[i, j])]
            ap = 0.0
            for k, ii, jj, kk in links:
                ap += kk
                if kk and not ident[ii, jj] and 1 <= ii <= nx and 1 <= jj <= ny:
                    coef[k, i, j] = -kk
            coef[0, i, j] = ap
    r = np.zeros((NX, NY))
    r[1:-1, 1:-1] = rng.standard_normal((nx, ny))
    r[ident] = 0.0
    return coef, r
