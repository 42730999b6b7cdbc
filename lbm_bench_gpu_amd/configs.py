"""The five benchmark workloads of this implementation.

Each workload is a periodic box of bubbles in a liquid with the ZSC physics
below.  Mapping of the workload wording onto the reference's inputs
(src/paramread.c:141-170):
  * tau_f is tauPhi (f is the phase distribution), tau_g is tauRho;
  * beta is the reference's alpha; kappa and beta follow from sigma and the
    interface width W;
  * phi* is given directly (1.0), overriding 0.5*(rho_H - rho_L); the mean
    density rho_mean stays 0.5*(rho_H + rho_L) = 5.5;
  * gravity is given directly (1e-6) and enters, as in the reference, as the
    +z buoyancy 2*phi*.g on phi > 0 cells;
  * "N steps" is N reference time steps (update, collide, stream).
Bubble centres come from the seeded generator (zsc_generate_bubbles) and are
kept R + W + 1 cells from every face, so the reference's non-periodic
initial profile and the periodic box agree.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .zsc import Inputs, derive_params, generate_bubbles


def physics(nx: int, ny: int, nz: int, **kw) -> Inputs:
    base = dict(xmax=nx, ymax=ny, zmax=nz, rho_l=1.0, rho_h=10.0, tau_rho=0.8, tau_phi=0.7,
                width=4.0, sigma=0.01, gamma=1.0, eo=0.0, phi_star=1.0, gravity=1e-6)
    base.update(kw)
    return Inputs(**base)


@dataclass
class Workload:
    name: str
    dims: tuple
    inputs: Inputs
    bubbles: np.ndarray
    steps: int
    precisions: tuple

    @property
    def cells(self) -> int:
        nx, ny, nz = self.dims
        return nx * ny * nz

    def params(self):
        return derive_params(self.inputs, float(self.bubbles[0, 3]) if len(self.bubbles) else 0.0)


def _gen(seed, n, rmin, rmax, dims, gap=4.0, width=4.0):
    return generate_bubbles(seed, n, rmin, rmax, gap, width, dims)


def workload(i: int) -> Workload:
    if i == 0:
        dims = (64, 64, 64)
        # centre in the reference's coordinates (x, y 0-based, z 1-based)
        b = np.array([[32.0, 32.0, 32.0, 10.0]])
        return Workload("zsc-64^3-1bubble", dims, physics(*dims), b, 1000, ("f64",))
    if i == 1:
        dims = (256, 256, 256)
        return Workload("zsc-256^3-16bubbles", dims, physics(*dims), _gen(1, 16, 10, 14, dims),
                        2000, ("f64", "f32"))
    if i == 2:
        dims = (512, 512, 512)
        return Workload("zsc-512^3-128bubbles", dims, physics(*dims), _gen(2, 128, 8, 16, dims),
                        1000, ("f32", "bf16", "fp16"))
    if i == 3:
        dims = (1024, 1024, 1024)
        return Workload("zsc-1024^3-512bubbles", dims, physics(*dims),
                        _gen(3, 512, 10, 20, dims), 500, ("bf16",))
    if i == 4:
        dims = (128, 128, 128)
        return Workload("zsc-128^3-8bubbles-adjoint", dims, physics(*dims),
                        _gen(4, 8, 12, 12, dims), 400, ("f32",))
    raise ValueError(f"no workload {i}")


def z_slab(i: int, parts: int, rank: int = 0) -> Workload:
    """Rank ``rank``'s z-slab of workload i cut into ``parts`` equal slabs
    (the z split of src/vgrid.c:37-47): the box nx x ny x nz/parts and the
    bubbles that lie wholly inside the slab (R + W from its faces), moved to
    its z origin.  A periodic run of it is the step at one GPU's share of the
    decomposed box, without the halo exchange that would join the slabs."""
    w = workload(i)
    nx, ny, nz = w.dims
    if parts < 1 or nz % parts or not 0 <= rank < parts:
        raise ValueError("parts must divide nz and 0 <= rank < parts")
    nzs = nz // parts
    z0 = rank * nzs
    width = w.inputs.width
    # z is 1-based: the slab holds planes z0 + 1 .. z0 + nzs
    keep = [b for b in w.bubbles
            if b[2] - b[3] - width >= z0 + 1 and b[2] + b[3] + width <= z0 + nzs]
    b = np.array(keep, dtype=np.float64).reshape(-1, 4)
    b[:, 2] -= z0
    dims = (nx, ny, nzs)
    inputs = physics(*dims, **{k: getattr(w.inputs, k) for k in
                              ("rho_l", "rho_h", "tau_rho", "tau_phi", "width", "sigma", "gamma",
                               "eo", "phi_star", "gravity")})
    return Workload(f"{w.name}-zslab{rank}of{parts}", dims, inputs, b, w.steps, w.precisions)
