"""Known-answer tests of the deviation codecs (bf16, fp16 and FP32 storage
of g as deviations) through the C-ABI: zsc_state_set encodes FP64 host values
into the context's lattice, zsc_state_get decodes them back.

The expected values are formed here, independently of the library:
  g_i (rest r_i = fl32(w_i) * fl32(rho_mean) in FP32; v - r_i and
  stored + r_i in FP64, where both are exact):
      bf16 / fp16: stored = RNE16(fl32(v - r_i)),  read = stored + r_i
      f32d:        stored = fl32(v - r_i),         read = stored + r_i
  f_1..6 (rest 0):  stored = RNE16(fl32(v)) (16-bit modes), fl32(v) (f32d)
  f_0 (FP32 in every mode):  read = fl32(v)
RNE16 is round-to-nearest-even to bf16 (torch's float32 -> bfloat16
conversion) or to fp16 (NumPy's float32 -> float16 conversion), and the
tie and subnormal vectors below also carry their rounded value in closed
form, so that a truncating or ties-away conversion fails them."""
import numpy as np
import pytest

from lbm_bench_gpu_amd import zsc
from tests.test_gpu_parity import small_case

pytestmark = pytest.mark.gpu

W = [1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12
F32 = np.float32


def bf16_rne(x):
    import torch
    return torch.from_numpy(np.asarray(x, dtype=F32)).to(torch.bfloat16).to(torch.float32).numpy()


def fp16_rne(x):
    with np.errstate(over="ignore"):
        return np.asarray(x, dtype=F32).astype(np.float16).astype(F32)


def tie_cases(prec):
    """Deviations exactly halfway between two neighbouring 16-bit values,
    (k + 1/2) u with u the spacing in [2^-10, 2^-9) for bf16 (u = 2^-17) and
    in [2^-8, 2^-7) for fp16 (u = 2^-18): RNE goes to the even k, where
    ties-away would not; plus values 2^-22 above (away from zero) and below
    (towards zero) each tie.  Every one is exact in FP32 next to any w_i rho."""
    u, ks = (2.0 ** -17, (128, 129, 200, 201, 254, 255)) if prec == "bf16" else \
        (2.0 ** -18, (1024, 1025, 1500, 1501, 2046, 2047))
    d, want = [], []
    for k in ks:
        for s in (1.0, -1.0):
            d.append(s * (k + 0.5) * u)
            want.append(s * (k + (k & 1)) * u)
            d.append(s * ((k + 0.5) * u + 2.0 ** -22))
            want.append(s * (k + 1) * u)
            d.append(s * ((k + 0.5) * u - 2.0 ** -22))
            want.append(s * k * u)
    return np.array(d), np.array(want)


def fp16_subnormal_cases():
    """fp16 subnormals are multiples of 2^-24: exact ones stay, ties
    (j + 1/2) 2^-24 go to the even j, and values under half the smallest
    one flush to zero; then a value past the largest finite fp16 (inf)."""
    d, want = [], []
    for j in (1, 2, 3, 7, 100, 1023):
        d += [j * 2.0 ** -24, (j + 0.5) * 2.0 ** -24]
        want += [j * 2.0 ** -24, (j + (j & 1)) * 2.0 ** -24]
    d += [0.25 * 2.0 ** -24, -0.5 * 2.0 ** -24, 65504.0, 70000.0, 1e-40]
    want += [0.0, 0.0, 65504.0, np.inf, 0.0]
    return np.array(d), np.array(want)


def rest_of(i, rho_mean):
    return np.float64(F32(W[i]) * F32(rho_mean))


def expected_round_trip(prec, f, g, rho_mean):
    fe = np.empty_like(f)
    ge = np.empty_like(g)
    fe[0] = f[0].astype(F32)
    if prec == "f32d":
        fe[1:] = f[1:].astype(F32)
        enc = lambda d: d
    else:
        fe[1:] = fp16_rne(f[1:].astype(F32)).astype(np.float64)
        enc = bf16_rne if prec == "bf16" else fp16_rne
    with np.errstate(invalid="ignore", over="ignore"):
        for i in range(19):
            rest = rest_of(i, rho_mean)
            dev = (g[i] - rest).astype(F32)
            ge[i] = enc(dev).astype(np.float64) + rest
    return fe, ge


def compare(prec, f, g, fe, ge, fo, go):
    bad_f = np.argwhere(~((fo == fe) | (np.isnan(fo) & np.isnan(fe))))
    bad_g = np.argwhere(~((go == ge) | (np.isnan(go) & np.isnan(ge))))
    print(prec, "codec mismatches f", len(bad_f), "g", len(bad_g), flush=True)
    assert len(bad_f) == 0, [(tuple(k), f[tuple(k)], fo[tuple(k)], fe[tuple(k)]) for k in bad_f[:5]]
    assert len(bad_g) == 0, [(tuple(k), g[tuple(k)], go[tuple(k)], ge[tuple(k)]) for k in bad_g[:5]]


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_codec_known_answers(need_gpu, prec):
    dims = (16, 8, 4)
    nx, ny, nz = dims
    n = nx * ny * nz
    p, phi0 = small_case(dims, seed=41)
    rng = np.random.default_rng(42)
    # background: random deviations of the size a run sees
    f = rng.normal(0.0, 1e-3, (7, n))
    f[0] = rng.uniform(-1.2, 1.2, n)
    g = np.array([W[i] * p.rho_mean + rng.normal(0.0, 2e-3, n) for i in range(19)])
    rests = [float(F32(W[i]) * F32(p.rho_mean)) for i in range(19)]
    # known answers, each placed in every g (resp. f_1..6) population
    dg, wg = tie_cases(prec)
    df, wf = fp16_subnormal_cases()
    for i in range(19):
        g[i, : dg.size] = rests[i] + dg   # exact in FP32: fl32(v) - r_i = d
    for i in range(1, 7):
        f[i, : df.size] = df
    specials = [np.nan, np.inf, -np.inf]
    g[5, dg.size: dg.size + 3] = specials
    f[3, df.size: df.size + 3] = specials
    f[0, :3] = [1e-40, -3.5e-42, np.nan]   # FP32 subnormals and NaN stored as they are
    fe, ge = expected_round_trip(prec, f, g, p.rho_mean)
    # the closed forms of the tie / subnormal vectors agree with the
    # expectation built from the library-independent converters
    for i in range(19):
        assert np.array_equal(ge[i, : dg.size], rests[i] + wg), i
    assert np.array_equal(fe[1, : df.size], wf.astype(F32).astype(np.float64))

    with zsc.Context(*dims, precision=prec) as ctx:
        ctx.set_params(p)
        ctx.state_init(phi0)
        ctx.state_set(f.reshape(7, nz, ny, nx), g.reshape(19, nz, ny, nx))
        fo, go = ctx.state_get()
    compare(prec, f, g, fe, ge, fo.reshape(7, n), go.reshape(19, n))


def test_f32d_keeps_the_deviation_bits(need_gpu):
    """FP32 storage of g as deviations at rho_mean = 500.5 (the reference's
    Input/properties.in, density ratio 1000): values rest + d with |d| from
    1e-12 to 1e-2 read back as rest + fl32(d) exactly, i.e. within 2^-24 |d|
    of the value written, where an FP32 store of the value itself rounds at
    ulp(w_i * 500.5) (up to 2^-18); f, NaN and +-Inf as in the other codecs."""
    dims = (16, 8, 4)
    nx, ny, nz = dims
    n = nx * ny * nz
    p, phi0 = small_case(dims, seed=43)
    p = p.replace(rho_mean=500.5)
    rng = np.random.default_rng(44)
    f = rng.normal(0.0, 1e-3, (7, n))
    f[0] = rng.uniform(-499.5, 499.5, n)
    d = rng.choice([-1.0, 1.0], (19, n)) * 10.0 ** rng.uniform(-12, -2, (19, n))
    g = np.array([rest_of(i, 500.5) + d[i] for i in range(19)])
    g[4, :3] = [np.nan, np.inf, -np.inf]
    fe, ge = expected_round_trip("f32d", f, g, 500.5)
    with zsc.Context(*dims, precision="f32d") as ctx:
        ctx.set_params(p)
        ctx.state_init(phi0)
        ctx.state_set(f.reshape(7, nz, ny, nx), g.reshape(19, nz, ny, nx))
        fo, go = ctx.state_get()
    fo, go = fo.reshape(7, n), go.reshape(19, n)
    compare("f32d", f, g, fe, ge, fo, go)
    ok = np.isfinite(g)
    assert np.all(np.abs(go - g)[ok] <= 2.0 ** -24 * np.abs(d)[ok])
    raw = np.array([g[i].astype(F32).astype(np.float64) for i in range(19)])
    print("f32d max |read - written| / |d|", float(np.max((np.abs(go - g) / np.abs(d))[ok])),
          "an FP32 store of the value:", float(np.max((np.abs(raw - g) / np.abs(d))[ok])), flush=True)
