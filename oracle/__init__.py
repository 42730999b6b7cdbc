"""FP64 CPU oracle for the ZSC bubbly-flow step -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP product path of ``lbm_bench_gpu_amd``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import it; the product never does.

It wraps ``liboracle.so`` (built from ``zsc_oracle.c`` by ``oracle/Makefile``),
an FP64 C restatement of the reference kernels (citations in the C source), and
adds NumPy helpers.  Arrays are NumPy float64, populations pop-major with shape
``(q, nz, ny, nx)`` -- flat index ``x + nx*(y + ny*z)`` per population.

Pinning of each function (the checks live in ``tests/test_oracle.py``):
  zo_derive           ZSC identities sigma = (4/3) sqrt(2 kappa beta) phi*^3 and
                      W = sqrt(2 kappa / beta) / phi*, plus the D3Q19 weights.
  zo_init_phi         tanh profile values, last-bubble-wins, 1-based z.
  zo_init_f/_g        the moments of the ZSC equilibria at u = 0 on a random
                      phi (mu != 0): sum f = phi, sum c f = 0,
                      sum c_a c_b f = Gamma mu d_ab; sum g = rho, sum c g = 0,
                      sum c_a c_b g = (rho/3 + phi mu) d_ab; f_1 = Gamma mu / 2
                      recovers the Laplacian's Fourier symbol.
  zo_update_*         brute-force NumPy sums on random states.
  zo_collision_f/_g   stencil symbols on Fourier modes, sum f_B = phi,
                      sum g_B = rho, the momentum closed form of the reference's
                      forcing; the first and second moments of the f collision
                      (C_f = phi invEta2 / tau_phi, A_f = Gamma mu / (2 tau_phi));
                      the advection speed eta / (1 - eta) u of a phi mode in a
                      uniform flow; fixed points at phi = +-phi* and at uniform
                      phi with mu != 0 (the phi*mu terms of collision_g).
                      Each of these checks fails on a planted mistake in the
                      term it pins (zo_set_mutation).
  zo_stream_f/_g      permutation/conservation, delta-input smoothing closed form.
  zo_run              an independent NumPy transcription that keeps the
                      reference's ghost planes, neighbour tables and pack/unpack
                      halo semantics (tests/ghost_ref.py), plus translation
                      covariance and exact conservation.
  zo_stats            brute-force NumPy sums.
  zo_pressure         bulk value rho c_s^2 at phi = +-phi*; double-well and
                      square-gradient terms on Fourier modes of phi (closed
                      forms from the stencil symbols); the pressure sums of
                      zo_stats.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.POINTER(ctypes.c_double)


class ZoCoef(ctypes.Structure):
    _fields_ = [(n, _D) for n in (
        "phi_star", "phi_star_sq", "rho_mean", "alpha", "alpha4", "kappa", "gamma",
        "eta", "eta2", "inv_eta2", "inv_tau_phi", "inv_tau_phi_one", "inv_tau_rho_one",
        "k0", "k1", "k2", "kc0", "kc1", "kc2", "gravity", "grav", "width", "sigma")]


def build() -> None:
    """Compile liboracle.so with make (called by the repo's build step)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.zo_derive.argtypes = [_D] * 11 + [ctypes.POINTER(ZoCoef)]
        L.zo_init_phi.argtypes = [_I, _I, _I, _P, _I, _D, _D, _P]
        L.zo_init_f.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), _P, _P]
        L.zo_init_g.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), _P, _P, _P]
        L.zo_update_phi.argtypes = [_I, _I, _I, _P, _P]
        L.zo_update_rho.argtypes = [_I, _I, _I, _P, _P]
        L.zo_update_velocity.argtypes = [_I, _I, _I, _P, _P, _P, _P, _P]
        L.zo_collision_f.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), _I] + [_P] * 7
        L.zo_collision_g.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), _I] + [_P] * 7
        L.zo_stream_f.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), _P, _P]
        L.zo_stream_g.argtypes = [_I, _I, _I, _P, _P]
        L.zo_run.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), _I, _I, _P, _P, _P]
        L.zo_stats.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef)] + [_P] * 6
        L.zo_pressure.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef)] + [_P] * 3
        L.zo_set_mutation.argtypes = [_I]
        for f in ("zo_derive", "zo_init_phi", "zo_init_f", "zo_init_g", "zo_update_phi",
                  "zo_update_rho", "zo_update_velocity", "zo_collision_f", "zo_collision_g",
                  "zo_stream_f", "zo_stream_g", "zo_run", "zo_stats", "zo_pressure",
                  "zo_set_mutation"):
            getattr(L, f).restype = None
        _lib = L
    return _lib


# Planted mistakes (zsc_oracle.h): each changes one term of one function.
MUT_NONE = 0
MUT_AF_NO_TAU = 1
MUT_CF_NO_TAU = 2
MUT_CF_HALF = 3
MUT_F0_ONE_AF = 4
MUT_INITG_REST_3PHIMU = 5
MUT_INITG_AXIS_PHIMU = 6
MUT_INITF_AF_NO_HALF = 7
MUT_CG_REST_3PHIMU = 8
MUT_CG_DIAG_6PHIMU = 9


class mutation:
    """Context manager that plants one mistake in the oracle for the duration
    of a block (sensitivity tests of the pins), then restores the reference's
    arithmetic."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        lib().zo_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().zo_set_mutation(MUT_NONE)
        return False


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


@dataclass
class Physics:
    """Raw inputs of Input/properties.in plus the overrides the configs need."""
    rho_l: float = 1.0
    rho_h: float = 10.0
    tau_rho: float = 0.8
    tau_phi: float = 0.7
    width: float = 4.0
    sigma: float = 0.01
    gamma: float = 1.0
    eo: float = 0.0
    r0: float = 10.0
    phi_star: float = 1.0      # <= 0: derive 0.5*(rho_h - rho_l)
    gravity: float = 1e-6      # < 0: derive from Eo


def derive(ph: Physics) -> ZoCoef:
    c = ZoCoef()
    lib().zo_derive(ph.rho_l, ph.rho_h, ph.tau_rho, ph.tau_phi, ph.width, ph.sigma,
                    ph.gamma, ph.eo, ph.r0, ph.phi_star, ph.gravity, ctypes.byref(c))
    return c


def init_phi(shape, bubbles, phi_star, width) -> np.ndarray:
    nz, ny, nx = shape
    b = np.ascontiguousarray(np.asarray(bubbles, dtype=np.float64).reshape(-1, 4))
    phi = np.empty((nz, ny, nx))
    lib().zo_init_phi(nx, ny, nz, _p(b), b.shape[0], phi_star, width, _p(phi))
    return phi


class State:
    """A lattice (post-stream, reference convention) of one periodic box."""

    def __init__(self, shape, coef: ZoCoef):
        self.nz, self.ny, self.nx = shape
        self.c = coef
        self.f = np.zeros((7,) + tuple(shape))
        self.g = np.zeros((19,) + tuple(shape))
        n = self.nx * self.ny * self.nz
        self._work = np.empty(31 * n)

    @property
    def dims(self):
        return self.nx, self.ny, self.nz

    def init(self, phi: np.ndarray, rho: np.ndarray | None = None):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if rho is None:
            rho = np.full_like(phi, self.c.rho_mean)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        lib().zo_init_f(*self.dims, ctypes.byref(self.c), _p(phi), _p(self.f))
        lib().zo_init_g(*self.dims, ctypes.byref(self.c), _p(phi), _p(rho), _p(self.g))
        return self

    def run(self, nsteps: int, grav_on: bool = True):
        lib().zo_run(*self.dims, ctypes.byref(self.c), int(grav_on), int(nsteps),
                     _p(self.f), _p(self.g), _p(self._work))
        return self

    def moments(self):
        return moments(self.f, self.g)


def moments(f: np.ndarray, g: np.ndarray):
    """phi, rho, ux, uy, uz of an A lattice (src/update_*.cu)."""
    shape = f.shape[1:]
    nz, ny, nx = shape
    phi, rho = np.empty(shape), np.empty(shape)
    ux, uy, uz = np.empty(shape), np.empty(shape), np.empty(shape)
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g)
    L = lib()
    L.zo_update_phi(nx, ny, nz, _p(f), _p(phi))
    L.zo_update_rho(nx, ny, nz, _p(g), _p(rho))
    L.zo_update_velocity(nx, ny, nz, _p(g), _p(rho), _p(ux), _p(uy), _p(uz))
    return phi, rho, ux, uy, uz


def collide(coef: ZoCoef, f, g, grav_on=True):
    """update_* then collision_f/_g: A lattice -> B lattice."""
    nz, ny, nx = f.shape[1:]
    phi, rho, ux, uy, uz = moments(f, g)
    fb, gb = np.empty_like(f), np.empty_like(g)
    L = lib()
    L.zo_collision_f(nx, ny, nz, ctypes.byref(coef), int(grav_on), _p(phi), _p(rho),
                     _p(ux), _p(uy), _p(uz), _p(np.ascontiguousarray(f)), _p(fb))
    L.zo_collision_g(nx, ny, nz, ctypes.byref(coef), int(grav_on), _p(phi), _p(rho),
                     _p(ux), _p(uy), _p(uz), _p(np.ascontiguousarray(g)), _p(gb))
    return fb, gb


def collide_with(coef: ZoCoef, phi, rho, ux, uy, uz, f, g, grav_on=True):
    """collision_f/_g from explicit moment fields."""
    nz, ny, nx = f.shape[1:]
    fb, gb = np.empty_like(f), np.empty_like(g)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi, rho, ux, uy, uz)]
    L = lib()
    L.zo_collision_f(nx, ny, nz, ctypes.byref(coef), int(grav_on), *map(_p, args),
                     _p(np.ascontiguousarray(f)), _p(fb))
    L.zo_collision_g(nx, ny, nz, ctypes.byref(coef), int(grav_on), *map(_p, args),
                     _p(np.ascontiguousarray(g)), _p(gb))
    return fb, gb


def stream(coef: ZoCoef, fb, gb):
    nz, ny, nx = fb.shape[1:]
    fa, ga = np.empty_like(fb), np.empty_like(gb)
    L = lib()
    L.zo_stream_f(nx, ny, nz, ctypes.byref(coef), _p(np.ascontiguousarray(fb)), _p(fa))
    L.zo_stream_g(nx, ny, nz, _p(np.ascontiguousarray(gb)), _p(ga))
    return fa, ga


def pressure(coef: ZoCoef, phi, rho) -> np.ndarray:
    """Pressure field (src/vtksave.c:75-119), FP64, shape of phi."""
    nz, ny, nx = phi.shape
    out = np.empty_like(phi, dtype=np.float64)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi, rho)]
    lib().zo_pressure(nx, ny, nz, ctypes.byref(coef), *map(_p, args), _p(out))
    return out


def stats(coef: ZoCoef, phi, rho, ux, uy, uz) -> np.ndarray:
    nz, ny, nx = phi.shape
    out = np.zeros(10)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi, rho, ux, uy, uz)]
    lib().zo_stats(nx, ny, nz, ctypes.byref(coef), *map(_p, args), _p(out))
    return out
