"""Error handling of the C-ABI on a live device: calls out of order, at the
wrong precision, with bad arguments, and a diverged state.  Every call
returns a status (no abort) and leaves the context usable."""
import ctypes

import numpy as np
import pytest

from lbm_bench_gpu_amd import zsc
from lbm_bench_gpu_amd._lib import (ZSC_ERR_ARG, ZSC_ERR_DIVERGED, ZSC_ERR_PRECISION,
                                    ZSC_ERR_STATE, ZscError, lib)
from tests.test_gpu_parity import small_case

pytestmark = pytest.mark.gpu


def _code(fn, *a):
    with pytest.raises(ZscError) as e:
        fn(*a)
    return e.value.code


def test_calls_out_of_order(need_gpu):
    dims = (16, 8, 6)
    p, phi0 = small_case(dims, seed=31)
    with zsc.Context(*dims, precision="f32") as ctx:
        assert _code(ctx.state_init, phi0) == ZSC_ERR_STATE      # no parameters yet
        ctx.set_params(p)
        assert _code(ctx.step, 1) == ZSC_ERR_STATE                # no state yet
        assert _code(ctx.stats) == ZSC_ERR_STATE
        assert _code(ctx.state_get) == ZSC_ERR_STATE
        assert _code(ctx.step, -1) == ZSC_ERR_ARG
        ctx.state_init(phi0)
        ctx.step(2)                                               # still usable
        assert np.isfinite(ctx.stats()["sum_phi"])


def test_dropin_and_adjoint_need_f32_or_f64(need_gpu):
    import torch
    dims = (8, 8, 4)
    n = 8 * 8 * 4
    p, _ = small_case(dims, seed=32)
    f = [torch.zeros(2 * n, dtype=torch.float32, device="cuda") for _ in range(7)]
    phi = torch.zeros(n, dtype=torch.float32, device="cuda")
    for prec in ("bf16", "fp16", "f32d"):
        with zsc.Context(*dims, precision=prec) as ctx:
            ctx.set_params(p)
            assert _code(ctx.init_f, phi, f) == ZSC_ERR_PRECISION
            assert _code(ctx.update_phi, phi, f) == ZSC_ERR_PRECISION
            if prec == "f32d":
                continue  # the adjoint runs over f32d states (tests/test_gpu_adjoint.py)
            buf = torch.zeros(26 * n, dtype=torch.float32, device="cuda")
            rc = lib().zsc_adj_init(ctx._h, ctypes.c_void_p(phi.data_ptr()),
                                    ctypes.c_void_p(buf.data_ptr()))
            assert rc == ZSC_ERR_PRECISION


def test_bad_launch_shape_is_refused(need_gpu):
    with zsc.Context(8, 8, 8, precision="f32") as ctx:
        assert _code(ctx.set_fused_config, 3, 16, 1) == ZSC_ERR_ARG   # ty not 1, 2, 4, 8
        assert _code(ctx.set_fused_config, 4, -1, 1) == ZSC_ERR_ARG
        assert _code(ctx.set_fused_config, 4, 16, 200 << 8) == ZSC_ERR_ARG  # LDS pad > 128 KiB


@pytest.mark.parametrize("prec", ["f64", "f32", "f32d", "bf16"])
def test_diverged_state_is_reported(need_gpu, prec):
    dims = (8, 6, 4)
    p, phi0 = small_case(dims, seed=33)
    with zsc.Context(*dims, precision=prec) as ctx:
        ctx.set_params(p)
        ctx.state_init(phi0)
        f, g = ctx.state_get()
        g[3, 1, 2, 3] = np.nan
        ctx.state_set(f, g)
        assert _code(ctx.stats) == ZSC_ERR_DIVERGED
        ctx.state_init(phi0)                                      # recovers
        ctx.step(1)
        assert np.isfinite(ctx.stats()["sum_rho"])


@pytest.mark.parametrize("prec,dims", [("f32", (64, 16, 8)), ("bf16", (128, 8, 8)),
                                       ("fp16", (63, 5, 3)), ("f64", (16, 8, 4)),
                                       ("f32d", (33, 5, 3))])
def test_copy_ceiling(need_gpu, prec, dims):
    """zsc_copy_ceiling times both copies on the context's lattices (the
    march copy only where the x extent suits its lane words) and refuses to
    run once a state is set, which it would overwrite."""
    p, phi0 = small_case(dims, seed=36)
    with zsc.Context(*dims, precision=prec) as ctx:
        ctx.set_params(p)
        c = ctx.copy_ceiling()
        assert c["linear"]["ms"] > 0 and c["bytes_per_copy"] > 0
        assert (c["march"]["ms"] > 0) == (prec in ("f32", "f32d", "f64") or dims[0] % 2 == 0)
        ctx.state_init(phi0)
        assert _code(ctx.copy_ceiling) == ZSC_ERR_STATE
        ctx.step(1)
        assert np.isfinite(ctx.stats()["sum_phi"])


def test_place_lattices_preconditions(need_gpu):
    dims = (16, 8, 4)
    p, phi0 = small_case(dims, seed=34)
    with zsc.Context(*dims, precision="f32") as ctx:
        assert _code(ctx.place_lattices, 2) == ZSC_ERR_STATE          # no parameters yet
        ctx.set_params(p)
        assert _code(ctx.place_lattices, 0) == ZSC_ERR_ARG
        assert _code(ctx.place_lattices, 17) == ZSC_ERR_ARG
        ms = ctx.place_lattices(2)
        assert ms[0] > 0.0 and len(ms) == 2
        assert _code(ctx.place_lattices, 2) == ZSC_ERR_STATE          # lattices exist
        ctx.state_init(phi0)
        ctx.step(2)
    with zsc.Context(*dims, precision="f32") as ctx:
        ctx.set_params(p)
        ctx.state_init(phi0)
        assert _code(ctx.place_lattices, 2) == ZSC_ERR_STATE          # state already set


def test_restart_file_validation_and_caller_values(need_gpu, tmp_path):
    """zsc_state_save_ex / zsc_state_load_ex: caller values survive the round
    trip; a file whose parameters were edited to eta = 1.5 or a non-finite
    invEta2 is refused with ZSC_ERR_IO before the context's state is touched;
    a version-1 file (no caller values) still loads to the same state."""
    import struct
    from lbm_bench_gpu_amd._lib import ZSC_ERR_IO, ZscParams
    dims = (16, 8, 6)
    p, phi0 = small_case(dims, seed=35)
    path = str(tmp_path / "s.zsc")
    with zsc.Context(*dims, precision="f64") as ctx:
        ctx.set_params(p)
        ctx.state_init(phi0)
        ctx.step(3)
        ctx.save(path, step=3, values=[1.5, -2.25])
        f_ref, g_ref = ctx.state_get()
    raw = bytearray(open(path, "rb").read())
    assert struct.unpack_from("<i", raw, 8)[0] == 2                 # header version
    hdr_v1 = 56 + ctypes.sizeof(ZscParams)                            # fixed part
    hdr_v2 = hdr_v1 + 8 + 8 * 8                                       # + count, pad, 8 values
    with zsc.Context(*dims, precision="f64") as ctx:
        step, vals = ctx.load(path, with_values=True)
        assert step == 3 and vals.tolist() == [1.5, -2.25]
        f, g = ctx.state_get()
        assert np.array_equal(f, f_ref) and np.array_equal(g, g_ref)
        for field, value in (("eta", 1.5), ("inv_eta2", float("inf"))):
            bad = bytearray(raw)
            struct.pack_into("<d", bad, 56 + getattr(ZscParams, field).offset, value)
            bpath = str(tmp_path / f"bad_{field}.zsc")
            open(bpath, "wb").write(bad)
            assert _code(ctx.load, bpath) == ZSC_ERR_IO, field
            f, g = ctx.state_get()                                    # state untouched
            assert np.array_equal(f, f_ref) and np.array_equal(g, g_ref)
        v1 = bytearray(raw[:hdr_v1]) + raw[hdr_v2:]
        struct.pack_into("<i", v1, 8, 1)
        open(str(tmp_path / "v1.zsc"), "wb").write(v1)
    with zsc.Context(*dims, precision="f64") as ctx:
        step, vals = ctx.load(str(tmp_path / "v1.zsc"), with_values=True)
        assert step == 3 and vals.size == 0
        f, g = ctx.state_get()
        assert np.array_equal(f, f_ref) and np.array_equal(g, g_ref)
        ctx.step(2)
        assert np.isfinite(ctx.stats()["sum_phi"])
