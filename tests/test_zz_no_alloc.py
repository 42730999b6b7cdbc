"""zsc_moment_stats_noalloc (no device allocation; scratch in a
library-owned device array) against zsc_moment_stats (context-allocated
scratch): bit-for-bit equal results on two calls in a row, the second on a
smaller box, and both equal to the oracle's reductions of the same arrays."""
import numpy as np
import pytest

import oracle as O
from lbm_bench_gpu_amd import zsc
from tests.test_gpu_parity import STATS_NAMES, _dropin_tensors, oracle_coef, small_case

pytestmark = pytest.mark.gpu


def _moments(dims, prec, seed, steps):
    """Drop-in moment arrays of a small bubble run after `steps` fused steps."""
    import torch
    nx, ny, nz = dims
    n = nx * ny * nz
    p, phi0 = small_case(dims, seed=seed)
    dt = torch.float64 if prec == "f64" else torch.float32
    f, g, mom = _dropin_tensors(n, dt)
    ctx = zsc.Context(*dims, precision=prec)
    ctx.set_params(p)
    ctx.state_init(phi0)
    ctx.step(steps)
    ctx.state_export(f, g)
    phi, rho, ux, uy, uz = mom
    ctx.update_phi(phi, f)
    ctx.update_rho(rho, g)
    ctx.update_velocity(rho, ux, uy, uz, g)
    ctx.sync()
    return ctx, p, mom


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_stats_equal_allocating_entry_point(need_gpu, prec):
    for dims, seed in (((40, 24, 20), 51), ((12, 10, 8), 52)):   # second call: smaller box
        print("noalloc stats", prec, dims, flush=True)
        ctx, p, mom = _moments(dims, prec, seed, 6)
        with ctx:
            s_no = ctx.moment_stats(*mom, noalloc=True)
            s_al = ctx.moment_stats(*mom)
            s_no2 = ctx.moment_stats(*mom, noalloc=True)    # array keeps the last call's data
            host = [m.cpu().numpy().astype(np.float64).reshape(dims[::-1]) for m in mom]
        assert s_no == s_al and s_no2 == s_al
        for k, r in zip(STATS_NAMES, O.stats(oracle_coef(p), *host)):
            assert abs(s_no[k] - r) <= 1e-10 * max(1.0, abs(r)), (k, s_no[k], r)
