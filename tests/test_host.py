"""Host-side logic of libzsc (no GPU): input files, parameter derivation,
bubble generator, host initial state, error paths."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from lbm_bench_gpu_amd import _lib, configs, zsc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_IN = os.path.join(GOLD, "reference_input")


def test_reads_reference_input_files():
    """Input/properties.in and Input/discrete.in as shipped with the reference
    (copied verbatim as fixtures)."""
    inp = zsc.read_inputs(os.path.join(REF_IN, "properties.in"))
    assert (inp.max_step, inp.relax_step, inp.save, inp.stat) == (1000, 1000, 2000, 100)
    assert (inp.xmax, inp.ymax, inp.zmax) == (128, 128, 128)
    assert (inp.block_x, inp.block_y) == (32, 8)
    assert (inp.rho_l, inp.rho_h, inp.tau_rho, inp.tau_phi) == (1.0, 1000.0, 1.0, 0.7)
    assert (inp.width, inp.sigma, inp.gamma, inp.eo) == (4.0, 0.01, 400.0, 10.0)
    assert inp.phi_star == 0.0 and inp.gravity == -1.0   # no overrides present
    b = zsc.read_bubbles(os.path.join(REF_IN, "discrete.in"))
    assert b.tolist() == [[64.0, 64.0, 64.0, 10.0]]


def test_derivation_matches_oracle_for_reference_inputs():
    inp = zsc.read_inputs(os.path.join(REF_IN, "properties.in"))
    p = zsc.derive_params(inp, 10.0)
    c = O.derive(O.Physics(rho_l=inp.rho_l, rho_h=inp.rho_h, tau_rho=inp.tau_rho,
                           tau_phi=inp.tau_phi, width=inp.width, sigma=inp.sigma,
                           gamma=inp.gamma, eo=inp.eo, r0=10.0, phi_star=-1.0, gravity=-1.0))
    for k, _ in O.ZoCoef._fields_:
        assert p.values[k] == getattr(c, k), k
    assert p.phi_star == 499.5 and p.rho_mean == 500.5


def test_workload_physics_matches_oracle():
    for i in range(5):
        w = configs.workload(i)
        p = w.params()
        c = O.derive(O.Physics(r0=float(w.bubbles[0, 3])))
        for k, _ in O.ZoCoef._fields_:
            assert p.values[k] == getattr(c, k), (i, k)


def test_overrides_and_errors(tmp_path):
    src = open(os.path.join(REF_IN, "properties.in")).read()
    good = tmp_path / "p.in"
    good.write_text(src + "\nphiStar 1.0\ngravity 1e-6\n")
    inp = zsc.read_inputs(str(good))
    assert inp.phi_star == 1.0 and inp.gravity == 1e-6
    p = zsc.derive_params(inp, 10.0)
    assert p.grav == 2e-6 and p.phi_star == 1.0
    cases = {
        "missing": (None, _lib.ZSC_ERR_IO),
        "truncated": (src[: len(src) // 2], _lib.ZSC_ERR_IO),
        "bad_override": (src + "\nfoo 1\n", _lib.ZSC_ERR_IO),
        "no_value": (src + "\nphiStar\n", _lib.ZSC_ERR_IO),
        "tau_half": (src.replace("1.00\t0.7", "0.50\t0.7"), _lib.ZSC_ERR_ARG),
        "zero_size": (src.replace("128\t128\t128", "0\t128\t128"), _lib.ZSC_ERR_ARG),
        # integers outside the range of int (2^32 + 24 would wrap to 24)
        "size_wraps": (src.replace("128\t128\t128", "4294967320\t128\t128"), _lib.ZSC_ERR_IO),
        "size_erange": (src.replace("128\t128\t128", "99999999999999999999999\t128\t128"),
                        _lib.ZSC_ERR_IO),
        "steps_negative_wrap": (src.replace("1000", "-4294967295", 1), _lib.ZSC_ERR_IO),
    }
    for name, (text, code) in cases.items():
        path = tmp_path / f"{name}.in"
        if text is not None:
            path.write_text(text)
        with pytest.raises(_lib.ZscError) as e:
            zsc.read_inputs(str(path))
        assert e.value.code == code, name
    with pytest.raises(_lib.ZscError) as e:
        zsc.read_bubbles(str(tmp_path / "nope.in"))
    assert e.value.code == _lib.ZSC_ERR_IO
    # a bubble count beyond int is rejected before anything is sized from it
    big = tmp_path / "many.in"
    big.write_text("nBubbles\n4294967297\n\nxc\tyc\tzc\tRc\n1\t1\t1\t1\n")
    with pytest.raises(_lib.ZscError) as e:
        zsc.read_bubbles(str(big))
    assert e.value.code == _lib.ZSC_ERR_IO


def test_allocating_calls_retry_once_after_releasing_torch_cache(monkeypatch):
    """An entry point that fails with ZSC_ERR_ALLOC is called again once after
    torch's cached blocks are returned, and the error stands when there was
    nothing to return or the second call fails too."""
    calls = []

    def fake(rcs):
        def fn(*args):
            calls.append(args)
            return rcs[len(calls) - 1]
        return fn

    monkeypatch.setattr(zsc, "_release_torch_cache", lambda: True)
    assert zsc._alloc_call(fake([_lib.ZSC_ERR_ALLOC, _lib.ZSC_OK]), 1, 2) == _lib.ZSC_OK
    assert calls == [(1, 2), (1, 2)]
    calls.clear()
    with pytest.raises(_lib.ZscError) as e:
        zsc._alloc_call(fake([_lib.ZSC_ERR_ALLOC, _lib.ZSC_ERR_ALLOC]))
    assert e.value.code == _lib.ZSC_ERR_ALLOC and len(calls) == 2
    calls.clear()
    monkeypatch.setattr(zsc, "_release_torch_cache", lambda: False)
    with pytest.raises(_lib.ZscError):
        zsc._alloc_call(fake([_lib.ZSC_ERR_ALLOC, _lib.ZSC_OK]))
    assert len(calls) == 1
    calls.clear()
    monkeypatch.setattr(zsc, "_release_torch_cache", lambda: True)
    with pytest.raises(_lib.ZscError) as e:   # other codes are not retried
        zsc._alloc_call(fake([_lib.ZSC_ERR_ARG, _lib.ZSC_OK]))
    assert e.value.code == _lib.ZSC_ERR_ARG and len(calls) == 1


def test_bubble_file_round_trip(tmp_path):
    b = zsc.generate_bubbles(5, 6, 3.0, 5.0, 2.0, 4.0, (64, 48, 40))
    path = str(tmp_path / "d.in")
    zsc.write_bubbles(path, b)
    assert np.array_equal(zsc.read_bubbles(path), b)


@pytest.mark.parametrize("i", [1, 2, 3, 4])
def test_generated_bubbles_match_fixtures(i):
    w = configs.workload(i)
    gold = zsc.read_bubbles(os.path.join(GOLD, f"discrete_seed{i}.in"))
    assert np.array_equal(w.bubbles, gold)


@pytest.mark.parametrize("i", [1, 2, 3, 4])
def test_generated_bubbles_respect_constraints(i):
    w = configs.workload(i)
    nx, ny, nz = w.dims
    b = w.bubbles
    W = w.inputs.width
    rmin, rmax = {1: (10, 14), 2: (8, 16), 3: (10, 20), 4: (12, 12)}[i]
    assert ((b[:, 3] >= rmin) & (b[:, 3] <= rmax)).all()
    m = b[:, 3] + W + 1
    assert (b[:, 0] >= m - 1e-4).all() and (b[:, 0] <= nx - 1 - m + 1e-4).all()
    assert (b[:, 1] >= m - 1e-4).all() and (b[:, 1] <= ny - 1 - m + 1e-4).all()
    assert (b[:, 2] >= 1 + m - 1e-4).all() and (b[:, 2] <= nz - m + 1e-4).all()
    d = np.sqrt(((b[:, None, :3] - b[None, :, :3]) ** 2).sum(-1))
    need = b[:, None, 3] + b[None, :, 3] + 4.0
    np.fill_diagonal(d, np.inf)
    assert (d >= need - 1e-9).all()


def test_generator_is_deterministic_and_fails_cleanly():
    a = zsc.generate_bubbles(1, 5, 3, 4, 1, 2, (40, 40, 40))
    b = zsc.generate_bubbles(1, 5, 3, 4, 1, 2, (40, 40, 40))
    c = zsc.generate_bubbles(2, 5, 3, 4, 1, 2, (40, 40, 40))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    with pytest.raises(_lib.ZscError) as e:
        zsc.generate_bubbles(1, 500, 10, 10, 4, 4, (40, 40, 40), max_tries=2000)
    assert e.value.code == _lib.ZSC_ERR_ARG


@pytest.mark.parametrize("i", [0, 1])
def test_host_init_matches_oracle(i):
    w = configs.workload(i)
    p = w.params()
    nx, ny, nz = w.dims
    if i == 1:
        nx, ny, nz = 128, 128, 128   # keep the CPU test small: a sub-box of the bubbles
        bubbles = w.bubbles[:2] * [0.5, 0.5, 0.5, 1.0]
    else:
        bubbles = w.bubbles
    a = zsc.init_phi_host((nx, ny, nz), bubbles, p.phi_star, p.width)
    b = O.init_phi((nz, ny, nx), bubbles, p.phi_star, p.width)
    assert np.array_equal(a, b)


def test_bytes_per_cell():
    assert zsc.bytes_per_cell("f64") == 416
    assert zsc.bytes_per_cell("f32") == 208 and zsc.bytes_per_cell("f32d") == 208
    assert zsc.bytes_per_cell("bf16") == 108 and zsc.bytes_per_cell("fp16") == 108


def test_driver_reports_missing_inputs(tmp_path):
    exe = os.path.join(os.path.dirname(HERE), "lbm_bench_gpu_amd", "bin", "zsc_run")
    if not os.path.exists(exe):
        pytest.skip("zsc_run not built")
    r = subprocess.run([exe, "-d", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = subprocess.run([exe, "-p", "f8"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    for prec in ("f64", "f32", "f32d", "bf16", "fp16"):  # accepted: fails on the inputs
        r = subprocess.run([exe, "-p", prec, "-d", str(tmp_path)], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 1 and "cannot open" in r.stderr, prec


def test_malformed_input_files_give_status_codes(tmp_path):
    """Random, truncated and garbled properties.in / discrete.in files: the
    readers return ZSC_OK, ZSC_ERR_IO or ZSC_ERR_ARG and never crash (the same
    inputs, 3000 of them, also run clean under AddressSanitizer on a
    host-instrumented build)."""
    import ctypes
    import random
    lib = _lib.lib()
    good = open(os.path.join(REF_IN, "properties.in")).read()
    rng = random.Random(5)
    tok = ["0", "1", "-5", "1e308", "nan", "inf", "abc", "1.5e", "--", "9" * 30, " ", "\n", "#", "0.7"]
    path = str(tmp_path / "fuzz.in")
    inp = _lib.ZscInputs()
    n_out = ctypes.c_int()
    buf = (ctypes.c_double * 400)()
    allowed = {_lib.ZSC_OK, _lib.ZSC_ERR_IO, _lib.ZSC_ERR_ARG}
    for _ in range(300):
        kind = rng.randrange(4)
        if kind == 0:
            text = " ".join(rng.choice(tok) for _ in range(rng.randrange(0, 40)))
        elif kind == 1:
            text = "".join(chr(rng.randrange(1, 256)) for _ in range(rng.randrange(0, 400)))
        elif kind == 2:
            text = "\n".join(" ".join(f"{rng.uniform(-1e3, 1e3):.3f}" for _ in range(rng.randrange(0, 6)))
                             for _ in range(rng.randrange(0, 120)))
        else:
            text = good[: rng.randrange(0, len(good) + 1)] + rng.choice(tok) * rng.randrange(0, 50)
        open(path, "w", encoding="latin-1").write(text)
        assert lib.zsc_read_inputs(path.encode(), ctypes.byref(inp)) in allowed
        cap = rng.choice([0, 1, 5, 100])
        assert lib.zsc_read_bubbles(path.encode(), buf if cap else None, cap, ctypes.byref(n_out)) in allowed
