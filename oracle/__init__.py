"""FP64 CPU oracle for the ZSC bubbly-flow step -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP product path of ``lbm_bench_gpu_amd``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import it; the product never does.

It wraps ``liboracle.so`` (built from ``zsc_oracle.c`` by ``oracle/Makefile``),
an FP64 C restatement of the reference kernels (citations in the C source), and
adds NumPy helpers.  Arrays are NumPy float64, populations pop-major with shape
``(q, nz, ny, nx)`` -- flat index ``x + nx*(y + ny*z)`` per population.
``liboracle_f32.so`` is the same source evaluated in FP32 (the reference's
kernels compute in float): ``State(..., dtype=np.float32)`` runs the step
with float32 arrays, FP32 arithmetic and the coefficients rounded to FP32;
it is pinned against the FP64 restatement (agreement to FP32 rounding).

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
  FP32 evaluation     within FP32 rounding of the FP64 restatement, and not
                      equal to it (tests/test_oracle.py).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_LIB32_PATH = os.path.join(_HERE, "liboracle_f32.so")
_lib = None
_lib32 = None

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


def _declare(L, P):
    """Argument types of the step functions, arrays of pointer type P."""
    L.zo_init_f.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), P, P]
    L.zo_init_g.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), P, P, P]
    L.zo_update_phi.argtypes = [_I, _I, _I, P, P]
    L.zo_update_rho.argtypes = [_I, _I, _I, P, P]
    L.zo_update_velocity.argtypes = [_I, _I, _I, P, P, P, P, P]
    L.zo_run.argtypes = [_I, _I, _I, ctypes.POINTER(ZoCoef), _I, _I, P, P, P]
    for f in ("zo_init_f", "zo_init_g", "zo_update_phi", "zo_update_rho", "zo_update_velocity",
              "zo_run"):
        getattr(L, f).restype = None


def lib32():
    """The FP32 evaluation of the restatement (float32 arrays)."""
    global _lib32
    if _lib32 is None:
        if not os.path.exists(_LIB32_PATH):
            build()
        L = ctypes.CDLL(_LIB32_PATH)
        _declare(L, ctypes.POINTER(ctypes.c_float))
        _lib32 = L
    return _lib32


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


def _p32(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _step_lib(dtype):
    """(library, pointer helper) of a lattice dtype: float64 or float32."""
    if np.dtype(dtype) == np.float64:
        return lib(), _p
    if np.dtype(dtype) == np.float32:
        return lib32(), _p32
    raise TypeError(f"oracle lattices are float64 or float32, not {dtype}")


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
    """A lattice (post-stream, reference convention) of one periodic box;
    dtype float64 (the FP64 oracle) or float32 (its FP32 evaluation)."""

    def __init__(self, shape, coef: ZoCoef, dtype=np.float64):
        self.nz, self.ny, self.nx = shape
        self.c = coef
        self.dtype = np.dtype(dtype)
        self._L, self._pp = _step_lib(self.dtype)
        self.f = np.zeros((7,) + tuple(shape), dtype=self.dtype)
        self.g = np.zeros((19,) + tuple(shape), dtype=self.dtype)
        n = self.nx * self.ny * self.nz
        self._work = np.empty(31 * n, dtype=self.dtype)

    @property
    def dims(self):
        return self.nx, self.ny, self.nz

    def init(self, phi: np.ndarray, rho: np.ndarray | None = None):
        L, pp = self._L, self._pp
        phi = np.ascontiguousarray(phi, dtype=self.dtype)
        if rho is None:
            rho = np.full_like(phi, self.c.rho_mean)
        rho = np.ascontiguousarray(rho, dtype=self.dtype)
        L.zo_init_f(*self.dims, ctypes.byref(self.c), pp(phi), pp(self.f))
        L.zo_init_g(*self.dims, ctypes.byref(self.c), pp(phi), pp(rho), pp(self.g))
        return self

    def run(self, nsteps: int, grav_on: bool = True):
        self._L.zo_run(*self.dims, ctypes.byref(self.c), int(grav_on), int(nsteps),
                       self._pp(self.f), self._pp(self.g), self._pp(self._work))
        return self

    def moments(self):
        return moments(self.f, self.g)


def moments(f: np.ndarray, g: np.ndarray):
    """phi, rho, ux, uy, uz of an A lattice (src/update_*.cu), at the
    lattice's precision."""
    shape = f.shape[1:]
    nz, ny, nx = shape
    L, pp = _step_lib(f.dtype)
    phi, rho = np.empty(shape, f.dtype), np.empty(shape, f.dtype)
    ux, uy, uz = np.empty(shape, f.dtype), np.empty(shape, f.dtype), np.empty(shape, f.dtype)
    f = np.ascontiguousarray(f)
    g = np.ascontiguousarray(g, dtype=f.dtype)
    L.zo_update_phi(nx, ny, nz, pp(f), pp(phi))
    L.zo_update_rho(nx, ny, nz, pp(g), pp(rho))
    L.zo_update_velocity(nx, ny, nz, pp(g), pp(rho), pp(ux), pp(uy), pp(uz))
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
