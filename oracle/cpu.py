"""ctypes wrapper of the C/OpenMP oracle (oracle/cpu_simple.c) -- test/baseline only."""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_cpu.so"


class OgGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("east", ctypes.c_int32), ("pin", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("rho", ctypes.c_double), ("mu", ctypes.c_double),
                ("dt", ctypes.c_double), ("alpha_u", ctypes.c_double), ("alpha_p", ctypes.c_double),
                ("alpha_m", ctypes.c_double), ("bc", (ctypes.c_double * 2) * 4)]


def build():
    """Compile liboracle_cpu.so with make (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle_cpu.so"], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        h.og_bicgstab.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_double, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int)]
        h.og_simple_step.argtypes = [ctypes.POINTER(OgGrid), P, P, P, P, P, P, P, P, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        h.og_num_threads.restype = ctypes.c_int
        _lib = h
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def grid(case) -> OgGrid:
    g = OgGrid()
    g.nx, g.ny, g.east, g.pin = case.nx, case.ny, int(case.east), int(bool(case.pin_pcor))
    g.dx, g.dy, g.rho, g.mu, g.dt = case.dx, case.dy, case.rho, case.mu, case.dt
    g.alpha_u, g.alpha_p, g.alpha_m = case.alpha_u, case.alpha_p, case.alpha_m
    for s, side in enumerate("wens"):
        g.bc[s][0], g.bc[s][1] = case.bc[side]
    return g


def bicgstab(coef: np.ndarray, b: np.ndarray, eps=1e-8, max_iter=0):
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    it = ctypes.c_int()
    st = lib().og_bicgstab(b.shape[0], b.shape[1], _p(coef), _p(b), _p(x), eps, max_iter, ctypes.byref(it))
    return x, {"status": st, "iterations": it.value}


class State:
    """Padded FP64 fields for og_simple_step."""

    def __init__(self, case, u, v, p, u0, v0, mask):
        self.case = case
        self.g = grid(case)
        self.u, self.v, self.p, self.u0, self.v0 = (np.array(a, dtype=np.float64, order="C") for a in (u, v, p, u0, v0))
        self.mask = np.ascontiguousarray(mask, dtype=np.uint8)
        self.pcor = np.zeros(case.shape_p)
        self.mdiv = np.zeros(case.shape_p)
        self.iters = (ctypes.c_int * 3)()

    def step(self, eps_mom=None, eps_p=None, max_iter=0):
        c = self.case
        st = lib().og_simple_step(ctypes.byref(self.g), _p(self.u), _p(self.v), _p(self.p), _p(self.u0), _p(self.v0),
                                  _p(self.mask), _p(self.pcor), _p(self.mdiv),
                                  c.eps_mom if eps_mom is None else eps_mom, c.eps_p if eps_p is None else eps_p,
                                  max_iter, self.iters)
        return st, list(self.iters)

    def time_forward(self):
        self.u0[...] = self.u
        self.v0[...] = self.v


def threads() -> int:
    return lib().og_num_threads()


def cores_available() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1
