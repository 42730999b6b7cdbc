"""The benchmark boxes at their full sizes, through properties that do not
depend on size.

* 256^3, 16 bubbles (seed 1): FP64 fused steps equal the FP64 oracle bit for
  bit, as on the small parity boxes.
* 512^3, 128 bubbles (seed 2), 1000 steps: FP32 compute over FP32, bf16,
  fp16 and f32d (FP32 deviations of g) storage against an FP64 run of the
  same box on the GPU (the FP64 path is bit-identical to the oracle, which
  would take about half an hour for this box on the host): max |dphi| within
  the stated bounds, 1e-3 for bf16, 1e-5 for f32d (measured 3.5e-6), and
  total phi and mass conserved to round-off (3e-6 relative: with FP32
  storage the bulk cells, all alike, round their over-relaxed g update the
  same way, so the mass error adds up to 1.6e-6 before the rounded map
  settles; FP64 conserves to 1e-15).
* 1024^3, 512 bubbles (seed 3): 100 steps of bf16 storage (the
  two-cells-per-lane kernel at its 32-bit offset limit) against FP32 storage
  of the same box, with total phi and mass conserved.
* 1024 x 1024 x 128, one GPU's z-slab of that box: FP32 and bf16 storage
  against FP64 over 200 steps, as at 512^3.
"""
import numpy as np
import pytest

import oracle as O
from lbm_bench_gpu_amd import configs, zsc

pytestmark = pytest.mark.gpu


def oracle_coef(p: zsc.Params):
    c = O.ZoCoef()
    for k, _ in O.ZoCoef._fields_:
        setattr(c, k, p.values[k])
    return c


def test_256_box_fp64_bitwise_vs_oracle(need_gpu):
    w = configs.workload(1)
    p = w.params()
    nx, ny, nz = w.dims
    phi0 = zsc.init_phi_host(w.dims, w.bubbles, p.phi_star, p.width)
    st = O.State((nz, ny, nx), oracle_coef(p)).init(phi0)
    st.run(2)
    with zsc.Context(*w.dims, precision="f64") as ctx:
        ctx.set_params(p)
        ctx.state_init(phi0)
        ctx.step(2)
        f, g = ctx.state_get()
    df = float(np.abs(f - st.f).max())
    dg = float(np.abs(g - st.g).max())
    print("256^3 f64 vs oracle: max|df|", df, "max|dg|", dg, flush=True)
    assert df == 0.0 and dg == 0.0


def test_256_box_fp32_2000_steps_vs_fp64(need_gpu):
    """The 256^3 box over its full 2000 steps: FP32 storage and compute
    against an FP64 run of the same box on the GPU (bit-identical to the
    oracle, as above).  Measured max|dphi| 4.6e-4, max|drho| 4.9e-5; bounds
    1e-3 and 1e-4.  Total phi drifts by 1.8e-6 and mass by -4.1e-6
    (relative; bounds 3e-6 and 1e-5): FP32 rounding of the over-relaxed g
    update, alike in the alike bulk cells, adds up over the 2000 steps where
    FP64 conserves to 1e-15."""
    w = configs.workload(1)
    p = w.params()
    out = {}
    for pr in ("f64", "f32"):
        with zsc.Context(*w.dims, precision=pr) as ctx:
            ctx.set_params(p)
            ctx.state_init_bubbles(w.bubbles)
            s0 = ctx.stats()
            ctx.step(2000)
            s1 = ctx.stats()
            phi, rho, *_ = ctx.moments()
            out[pr] = (phi, rho, s0, s1)
    e_phi = float(np.abs(out["f32"][0] - out["f64"][0]).max())
    e_rho = float(np.abs(out["f32"][1] - out["f64"][1]).max())
    _, _, s0, s1 = out["f32"]
    d_phi = (s1["sum_phi"] - s0["sum_phi"]) / abs(s0["sum_phi"])
    d_rho = (s1["sum_rho"] - s0["sum_rho"]) / abs(s0["sum_rho"])
    print(f"256^3 f32 x 2000: max|dphi| {e_phi:.3e} max|drho| {e_rho:.3e} "
          f"sum phi drift {d_phi:.2e} sum rho drift {d_rho:.2e}", flush=True)
    assert e_phi <= 1e-3 and e_rho <= 1e-4
    assert abs(d_phi) <= 3e-6 and abs(d_rho) <= 1e-5


STEPS_512 = 1000


@pytest.fixture(scope="module")
def ref_512(need_gpu):
    w = configs.workload(2)
    p = w.params()
    with zsc.Context(*w.dims, precision="f64") as ctx:
        ctx.set_params(p)
        ctx.state_init_bubbles(w.bubbles)
        s0 = ctx.stats()
        ctx.step(STEPS_512)
        s1 = ctx.stats()
        phi = ctx.moments()[0]
    return w, p, phi, s0, s1


@pytest.mark.parametrize("prec,tol_phi,tol_sum", [("f32", 1e-3, 3e-6), ("bf16", 1e-3, 3e-6),
                                                  ("fp16", 1e-4, 3e-6), ("f32d", 1e-5, 3e-6)])
def test_512_box_reduced_precision_vs_fp64(ref_512, prec, tol_phi, tol_sum):
    w, p, phi_ref, r0, r1 = ref_512
    with zsc.Context(*w.dims, precision=prec) as ctx:
        ctx.set_params(p)
        ctx.state_init_bubbles(w.bubbles)
        s0 = ctx.stats()
        ctx.step(STEPS_512)
        s1 = ctx.stats()
        phi = ctx.moments()[0]
    e_phi = float(np.abs(phi - phi_ref).max())
    d_phi = (s1["sum_phi"] - s0["sum_phi"]) / abs(s0["sum_phi"])
    d_rho = (s1["sum_rho"] - s0["sum_rho"]) / abs(s0["sum_rho"])
    print(f"512^3 {prec}: max|dphi| {e_phi:.3e}, sum phi drift {d_phi:.2e}, "
          f"sum rho drift {d_rho:.2e}; FP64 drift "
          f"{(r1['sum_phi'] - r0['sum_phi']) / abs(r0['sum_phi']):.2e}", flush=True)
    assert e_phi <= tol_phi
    assert abs(d_phi) <= tol_sum and abs(d_rho) <= tol_sum


def _phi_only(ctx, n):
    """phi of the context's state as an FP64 host array (the other moments
    not downloaded: at 1024^3 each is 8.6 GB)."""
    import ctypes
    from lbm_bench_gpu_amd._lib import check, lib
    phi = np.empty(n)
    check(lib().zsc_state_moments(ctx._h, phi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  None, None, None, None))
    return phi


def test_1024_box_bf16_vs_fp32(need_gpu):
    """1024^3, 512 bubbles, 100 steps: bf16 storage (the two-cells-per-lane
    kernel at the 32-bit offset limit of its FP32 f_0 pairs, 2^32 - 8 bytes;
    137 GB of lattices and work arrays) against FP32 storage of the same box
    (253 GB; an FP64 run would need 447 GB, so the FP64 comparison is the
    512^3 one above): max|dphi| within 1e-3 (measured 8.5e-5), and in both
    runs total phi and mass conserved to 1e-6 (relative; measured 3.3e-7 and
    -1.6e-7 at bf16, 1.6e-8 and -6.2e-7 at FP32)."""
    w = configs.workload(3)
    p = w.params()
    out = {}
    for prec in ("bf16", "f32"):
        print("1024^3", prec, flush=True)
        with zsc.Context(*w.dims, precision=prec) as ctx:
            ctx.set_params(p)
            ctx.state_init_bubbles(w.bubbles)
            s0 = ctx.stats()
            ctx.step(100)
            s1 = ctx.stats()
            out[prec] = _phi_only(ctx, w.cells)
        d_phi = (s1["sum_phi"] - s0["sum_phi"]) / abs(s0["sum_phi"])
        d_rho = (s1["sum_rho"] - s0["sum_rho"]) / abs(s0["sum_rho"])
        print(f"1024^3 {prec}: sum phi drift {d_phi:.2e}, sum rho drift {d_rho:.2e}, "
              f"volume ratio {s1['bubble_volume_full'] / s0['bubble_volume_full']:.6f}", flush=True)
        assert abs(d_phi) <= 1e-6 and abs(d_rho) <= 1e-6
    e_phi = float(np.abs(out["bf16"] - out["f32"]).max())
    print(f"1024^3 bf16 vs f32 after 100 steps: max|dphi| {e_phi:.3e}", flush=True)
    assert e_phi <= 1e-3


@pytest.mark.parametrize("prec,tol_phi", [("f32", 1e-3), ("bf16", 1e-3)])
def test_z_slab_share_vs_fp64(need_gpu, prec, tol_phi):
    """The 1024^3 box's z-slab for one GPU of eight (1024 x 1024 x 128, the
    bubbles wholly inside it): 200 steps at FP32 and bf16 storage against
    FP64 on the same GPU; total phi and mass conserved to round-off."""
    w = configs.z_slab(3, 8)
    p = w.params()
    out = {}
    for pr in ("f64", prec):
        with zsc.Context(*w.dims, precision=pr) as ctx:
            ctx.set_params(p)
            ctx.state_init_bubbles(w.bubbles)
            s0 = ctx.stats()
            ctx.step(200)
            s1 = ctx.stats()
            out[pr] = (ctx.moments()[0], s0, s1)
    phi, s0, s1 = out[prec]
    e_phi = float(np.abs(phi - out["f64"][0]).max())
    d_phi = (s1["sum_phi"] - s0["sum_phi"]) / abs(s0["sum_phi"])
    d_rho = (s1["sum_rho"] - s0["sum_rho"]) / abs(s0["sum_rho"])
    print(f"{w.name} {prec}: max|dphi| {e_phi:.3e}, sum phi drift {d_phi:.2e}, "
          f"sum rho drift {d_rho:.2e}", flush=True)
    assert len(w.bubbles) > 0
    assert e_phi <= tol_phi
    assert abs(d_phi) <= 3e-6 and abs(d_rho) <= 3e-6
