"""Host-side mirror of the reference's per-step interface over libzsc.

``Context`` exposes the nine reference kernels under their own names
(``init_f``, ``init_g``, ``update_phi``, ``update_rho``, ``update_velocity``,
``collision_f``, ``collision_g``, ``stream_f``, ``stream_g``; src/main.cu:344-566)
on device tensors laid out like the reference's arrays: one tensor per
population of ``2*N`` elements, A half first (src/main.cu:225-251).  It also
drives the fused fast path (``state_init`` / ``step`` / ``state_get``).
PyTorch only provides device memory and streams here; all arithmetic runs in
libzsc's HIP kernels.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import PARAM_FIELDS, STATS_FIELDS, ZscInputs, ZscParams, ZscStats, check, lib

_DP = ctypes.POINTER(ctypes.c_double)


def _dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "need a C-contiguous float64 array"
    return a.ctypes.data_as(_DP)


# ---------------------------------------------------------------- inputs ---
@dataclass
class Inputs:
    """Input/properties.in (src/paramread.c:40-59) plus the two overrides."""
    max_step: int = 1000
    relax_step: int = 0
    save: int = 1000000
    stat: int = 1000000
    xmax: int = 64
    ymax: int = 64
    zmax: int = 64
    block_x: int = 32
    block_y: int = 8
    rho_l: float = 1.0
    rho_h: float = 10.0
    tau_rho: float = 0.8
    tau_phi: float = 0.7
    width: float = 4.0
    sigma: float = 0.01
    gamma: float = 1.0
    eo: float = 0.0
    phi_star: float = 0.0      # > 0 overrides 0.5*(rho_h - rho_l)
    gravity: float = -1.0      # >= 0 overrides the Eo-derived gravity

    def to_c(self) -> ZscInputs:
        c = ZscInputs()
        for k, _ in ZscInputs._fields_:
            setattr(c, k, getattr(self, k))
        return c

    @classmethod
    def from_c(cls, c: ZscInputs) -> "Inputs":
        return cls(**{k: getattr(c, k) for k, _ in ZscInputs._fields_})


@dataclass
class Params:
    """Derived scalars (src/paramread.c:141-170), editable before set_params."""
    values: dict = field(default_factory=dict)

    def __getattr__(self, k):
        try:
            return self.values[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def to_c(self) -> ZscParams:
        c = ZscParams()
        for k in PARAM_FIELDS:
            setattr(c, k, float(self.values[k]))
        return c

    @classmethod
    def from_c(cls, c: ZscParams) -> "Params":
        return cls({k: getattr(c, k) for k in PARAM_FIELDS})

    def replace(self, **kw) -> "Params":
        v = dict(self.values)
        v.update(kw)
        return Params(v)


def read_inputs(path: str) -> Inputs:
    c = ZscInputs()
    check(lib().zsc_read_inputs(path.encode(), ctypes.byref(c)))
    return Inputs.from_c(c)


def read_bubbles(path: str) -> np.ndarray:
    n = ctypes.c_int(0)
    check(lib().zsc_read_bubbles(path.encode(), None, 0, ctypes.byref(n)))
    b = np.zeros((max(n.value, 1), 4))
    check(lib().zsc_read_bubbles(path.encode(), _dptr(b), n.value, ctypes.byref(n)))
    return b[: n.value].copy()


def write_bubbles(path: str, bubbles: np.ndarray) -> None:
    b = np.ascontiguousarray(bubbles, dtype=np.float64).reshape(-1, 4)
    check(lib().zsc_write_bubbles(path.encode(), _dptr(b), b.shape[0]))


def derive_params(inputs: Inputs, r0: float = 0.0) -> Params:
    c = ZscParams()
    check(lib().zsc_derive_params(ctypes.byref(inputs.to_c()), float(r0), ctypes.byref(c)))
    return Params.from_c(c)


def generate_bubbles(seed: int, n: int, r_min: float, r_max: float, gap: float, width: float,
                     dims, max_tries: int = 1_000_000) -> np.ndarray:
    nx, ny, nz = dims
    b = np.zeros((max(n, 1), 4))
    check(lib().zsc_generate_bubbles(seed, n, r_min, r_max, gap, width, nx, ny, nz, max_tries,
                                     _dptr(b)))
    return b[:n].copy()


def init_phi_host(dims, bubbles, phi_star: float, width: float) -> np.ndarray:
    nx, ny, nz = dims
    b = np.ascontiguousarray(bubbles, dtype=np.float64).reshape(-1, 4)
    phi = np.empty((nz, ny, nx))
    check(lib().zsc_init_phi_host(nx, ny, nz, _dptr(b), b.shape[0], phi_star, width, _dptr(phi)))
    return phi


def bytes_per_cell(precision: str) -> float:
    return lib().zsc_bytes_per_cell(_lib.PRECISIONS[precision])


# --------------------------------------------------------------- context ---
def _release_torch_cache() -> bool:
    """Returns torch's cached, unused device blocks to the driver; True when
    there were any."""
    import sys
    torch = sys.modules.get("torch")
    if torch is None or not torch.cuda.is_initialized():
        return False
    if torch.cuda.memory_reserved() <= torch.cuda.memory_allocated():
        return False
    torch.cuda.empty_cache()
    return True


def _alloc_call(fn, *args) -> int:
    """Calls a libzsc entry point that may allocate device memory; on
    ZSC_ERR_ALLOC, returns torch's cached blocks and tries once more.
    libzsc allocates with hipMalloc, and blocks that torch has freed but
    keeps cached are not available to it until torch returns them."""
    rc = fn(*args)
    if rc == _lib.ZSC_ERR_ALLOC and _release_torch_cache():
        rc = fn(*args)
    return check(rc)


class Context:
    """One periodic box on the current HIP device (zsc_create, which makes no
    device allocation; the lattices are allocated by the first call that
    needs them)."""

    def __init__(self, nx: int, ny: int, nz: int, precision: str = "f64", stream=None):
        self.dims = (nx, ny, nz)
        self.n = nx * ny * nz
        self.precision = precision
        self._h = ctypes.c_void_p()
        s = None if stream is None else ctypes.c_void_p(int(stream))
        check(lib().zsc_create(nx, ny, nz, _lib.PRECISIONS[precision], s, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().zsc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # parameters
    def set_params(self, p: Params):
        check(lib().zsc_set_params(self._h, ctypes.byref(p.to_c())))

    def set_gravity(self, on: bool):
        check(lib().zsc_set_gravity(self._h, int(bool(on))))

    def set_fused_config(self, ty: int, zchunk: int, flags: int = 0):
        check(lib().zsc_set_fused_config(self._h, ty, zchunk, flags))

    def fused_config(self) -> tuple:
        """(ty, zchunk, flags) of the fused kernel's launch shape in use."""
        v = [ctypes.c_int(0) for _ in range(3)]
        check(lib().zsc_get_fused_config(self._h, *map(ctypes.byref, v)))
        return tuple(x.value for x in v)

    def sync(self):
        check(lib().zsc_sync(self._h))

    # drop-in kernels on device tensors (torch tensors or raw pointers)
    @staticmethod
    def _ptr(t):
        return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())

    def _pops(self, ts, q):
        arr = (ctypes.c_void_p * q)()
        for i, t in enumerate(ts):
            arr[i] = None if t is None else self._ptr(t).value
        return arr

    def init_f(self, phi, f):
        check(lib().zsc_init_f(self._h, self._ptr(phi), self._pops(f, 7)))

    def init_g(self, phi, rho, g):
        check(lib().zsc_init_g(self._h, self._ptr(phi), self._ptr(rho), self._pops(g, 19)))

    def update_phi(self, phi, f):
        check(lib().zsc_update_phi(self._h, self._ptr(phi), self._pops(f, 7)))

    def update_rho(self, rho, g):
        check(lib().zsc_update_rho(self._h, self._ptr(rho), self._pops(g, 19)))

    def update_velocity(self, rho, ux, uy, uz, g):
        check(lib().zsc_update_velocity(self._h, self._ptr(rho), self._ptr(ux), self._ptr(uy),
                                        self._ptr(uz), self._pops(g, 19)))

    def collision_f(self, phi, rho, ux, uy, uz, f):
        check(lib().zsc_collision_f(self._h, *(self._ptr(t) for t in (phi, rho, ux, uy, uz)),
                                    self._pops(f, 7)))

    def collision_g(self, phi, rho, ux, uy, uz, g):
        check(lib().zsc_collision_g(self._h, *(self._ptr(t) for t in (phi, rho, ux, uy, uz)),
                                    self._pops(g, 19)))

    def stream_f(self, f):
        check(lib().zsc_stream_f(self._h, self._pops(f, 7)))

    def stream_g(self, g):
        check(lib().zsc_stream_g(self._h, self._pops(g, 19)))

    # fast path
    def state_init(self, phi: np.ndarray, rho: np.ndarray | None = None):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert phi.size == self.n
        rho = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        _alloc_call(lib().zsc_state_init, self._h, _dptr(phi), _dptr(rho))

    def state_init_bubbles(self, bubbles: np.ndarray):
        b = np.ascontiguousarray(bubbles, dtype=np.float64).reshape(-1, 4)
        _alloc_call(lib().zsc_state_init_bubbles, self._h, _dptr(b), b.shape[0])

    def step(self, nsteps: int):
        check(lib().zsc_step(self._h, int(nsteps)))

    def step_timed(self, nsteps: int) -> float:
        ms = ctypes.c_double(0.0)
        check(lib().zsc_step_timed(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def place_lattices(self, trials: int = 4) -> list:
        """zsc_place_lattices: keeps the fastest of up to `trials` lattice
        allocations; returns each trial's ms per step (0: not run)."""
        ms = np.zeros(int(trials))
        _alloc_call(lib().zsc_place_lattices, self._h, int(trials), _dptr(ms))
        return [float(x) for x in ms]

    def copy_ceiling(self) -> dict:
        """zsc_copy_ceiling on the context's own lattice pair (before the
        state is set): ms per copy and GB/s of read + write traffic for a
        linear copy and for the fused step's access pattern."""
        ms = np.zeros(2)
        _alloc_call(lib().zsc_copy_ceiling, self._h, _dptr(ms))
        lat = bytes_per_cell(self.precision) / 2 * self.n
        out = {"bytes_per_copy": 2 * lat}
        for name, t in zip(("linear", "march"), ms):
            out[name] = {"ms": round(float(t), 4),
                         "GB/s": round(2 * lat / (t * 1e-3) / 1e9, 1) if t > 0 else None}
        return out

    def state_get(self):
        nx, ny, nz = self.dims
        f = np.empty((7, nz, ny, nx))
        g = np.empty((19, nz, ny, nx))
        _alloc_call(lib().zsc_state_get, self._h, _dptr(f), _dptr(g))
        return f, g

    def state_set(self, f: np.ndarray, g: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        _alloc_call(lib().zsc_state_set, self._h, _dptr(f), _dptr(g))

    def moments(self):
        nx, ny, nz = self.dims
        out = [np.empty((nz, ny, nx)) for _ in range(5)]
        _alloc_call(lib().zsc_state_moments, self._h, *map(_dptr, out))
        return tuple(out)

    def moment_stats(self, phi, rho, ux, uy, uz, noalloc: bool = False) -> dict:
        """zsc_moment_stats: the diagnostics of src/stats.c / src/logsave.c
        over drop-in moment arrays (device tensors of N elements); with
        `noalloc`, zsc_moment_stats_noalloc, which allocates nothing."""
        s = ZscStats()
        args = (self._h, *(self._ptr(t) for t in (phi, rho, ux, uy, uz)), ctypes.byref(s))
        if noalloc:
            check(lib().zsc_moment_stats_noalloc(*args))
        else:
            _alloc_call(lib().zsc_moment_stats, *args)
        return {k: getattr(s, k) for k in STATS_FIELDS}

    def stats(self) -> dict:
        s = ZscStats()
        check(lib().zsc_state_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k in STATS_FIELDS}

    def pressure(self) -> np.ndarray:
        """Pressure of the A state (src/vtksave.c:75-119), FP64 (nz, ny, nx)."""
        nx, ny, nz = self.dims
        out = np.empty((nz, ny, nx))
        _alloc_call(lib().zsc_state_pressure, self._h, _dptr(out))
        return out

    def vtk_save(self, path: str, binary: bool = False):
        """Legacy VTK file of phi, pressure and velocity (src/vtksave.c)."""
        _alloc_call(lib().zsc_vtk_save, self._h, os.fsencode(path), 1 if binary else 0)

    def save(self, path: str, step: int = 0, values=()):
        """Restart file of the state as held, with up to eight caller values
        (zsc_state_save_ex); load() continues it bit for bit."""
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        check(lib().zsc_state_save_ex(self._h, os.fsencode(path), int(step),
                                      _dptr(v) if v.size else None, int(v.size)))

    def load(self, path: str, with_values: bool = False):
        """Restores state, parameters and gravity switch; returns the step,
        or (step, caller values) with `with_values`."""
        st = ctypes.c_int64(0)
        buf = np.zeros(_lib.RESTART_USER_MAX)
        nv = ctypes.c_int(0)
        _alloc_call(lib().zsc_state_load_ex, self._h, os.fsencode(path), ctypes.byref(st),
                    _dptr(buf), buf.size, ctypes.byref(nv))
        return (st.value, buf[: nv.value].copy()) if with_values else st.value

    def state_export(self, f, g):
        check(lib().zsc_state_export(self._h, self._pops(f, 7), self._pops(g, 19)))

    def state_import(self, f, g):
        _alloc_call(lib().zsc_state_import, self._h, self._pops(f, 7), self._pops(g, 19))
