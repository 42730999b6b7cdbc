"""Hand-written adjoint (zsc_adj_*) against independent derivatives:
  * per step: <J v, w> = <v, J^T w> with J v from central differences (FP64);
  * whole run: FP64 gradients against a taped FP64 CPU gradient (torch
    autograd over oracle/taped.py) on a 32^3 box;
  * the 128^3 adjoint workload: gradients against central finite differences
    (h = 1e-4 relative) of FP64 runs.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from lbm_bench_gpu_amd import adjoint, configs, zsc
from lbm_bench_gpu_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _runner_kw(precision):
    """AdjointRunner arguments of a test precision: "f64", "f32", "f32/f64"
    (FP32 adjoint lattices over FP64 states) or "f32/f32d" (over FP32 states
    whose g are deviations from w_i rho_mean)."""
    lam, _, states = precision.partition("/")
    return {"precision": lam, "states": states or lam}


def small_problem(dims, seed=0):
    nx, ny, nz = dims
    p = zsc.derive_params(configs.physics(*dims, gravity=1e-3), 4.0)
    rng = np.random.default_rng(seed)
    b = np.array([[nx / 2 + 0.3, ny / 2 - 0.2, nz / 2 + 1.1, 4.0], [2.0, 3.0, 2.5, 3.0]])
    phi0 = O.init_phi((nz, ny, nx), b, p.phi_star, p.width)
    phi0 = phi0 + 0.02 * rng.standard_normal(phi0.shape)
    return p, phi0


def test_step_dot_product(need_gpu):
    import torch
    dims = (12, 10, 8)
    p, phi0 = small_problem(dims)
    run = adjoint.AdjointRunner(dims, p, precision="f64", checkpoint_every=4)
    a = run.forward(adjoint.theta_of(p), torch.from_numpy(phi0), 3)
    n = run.n
    g = torch.Generator(device="cuda").manual_seed(1)
    v = torch.randn((26, n), dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn((26, n), dtype=torch.float64, device="cuda", generator=g)
    eps = 1e-6
    ap, am = a + eps * v, a - eps * v
    op, om = run._lattice(), run._lattice()
    lam = run._lattice()
    torch.cuda.synchronize()
    check(lib().zsc_adj_forward(run.ctx._h, _vp(ap), _vp(op), _vp(run.scratch)))
    check(lib().zsc_adj_forward(run.ctx._h, _vp(am), _vp(om), _vp(run.scratch)))
    gp = ctypes.cast(run.grads.data_ptr(), ctypes.POINTER(ctypes.c_double))
    check(lib().zsc_adj_backward(run.ctx._h, _vp(a), 0, None, _vp(w), _vp(lam), _vp(run.scratch),
                                 gp))
    run.ctx.sync()
    jv = (op - om) / (2 * eps)
    lhs = float((jv * w).sum())
    rhs = float((v * lam).sum())
    print("dot product", lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs)


@pytest.mark.parametrize("tape", [True, False])
def test_fp64_gradients_match_taped_cpu(need_gpu, tape):
    import torch
    from oracle import taped
    dims = (32, 32, 32)
    nx, ny, nz = dims
    p = zsc.derive_params(configs.physics(*dims), 6.0)
    b = np.array([[16.3, 15.6, 16.9, 6.0]])
    phi0 = O.init_phi((nz, ny, nx), b, p.phi_star, p.width)
    steps = 30
    theta = adjoint.theta_of(p)
    # target from perturbed physics, as the 128^3 workload
    tt = [theta[0] * 1.2, theta[1] * 0.9, theta[2], theta[3], theta[4] * 1.5]
    run = adjoint.AdjointRunner(dims, p, precision="f64", checkpoint_every=7, tape=tape)
    tgt = run.forward(tt, torch.from_numpy(phi0), steps, keep=False)[:7].sum(0).clone()
    L, grads = run.loss_and_grad(theta, torch.from_numpy(phi0), tgt, steps)
    assert (run.tape is not None) == tape
    # taped FP64 CPU reference (init inside the tape)
    th = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in theta]
    c = taped.coefficients(*th, phi_star=p.phi_star, gamma=p.gamma)
    f, g = taped.init_state(torch.from_numpy(phi0), c, p.rho_mean)
    for _ in range(steps):
        f, g = taped.step(f, g, c)
    tgt_cpu = tgt.view(nz, ny, nx).cpu()
    loss = ((f.sum(0) - tgt_cpu) ** 2).sum()
    ref = [float(x) for x in torch.autograd.grad(loss, th)]
    print("loss", L, float(loss), "\nadjoint", grads, "\ntaped  ", ref, flush=True)
    assert abs(L - float(loss)) <= 1e-10 * abs(float(loss))
    for a, r in zip(grads, ref):
        assert abs(a - r) <= 1e-8 * max(abs(r), 1e-30), (a, r)


@pytest.mark.parametrize("precision", ["f64", "f32", "f32/f64", "f32/f32d"])
def test_tape_equals_recomputation(need_gpu, precision):
    """The tape holds the same states and phi_n that the checkpointed
    backward pass recomputes (the same kernels on the same inputs), so both
    give the same gradients and the same dL/da_0 bit for bit; segments of 4
    with a ragged last one (23 steps: the last segment kept from the forward
    pass, five recomputed, so the overlapped pass reuses each of its two
    buffer sets)."""
    import torch
    dims = (14, 10, 8)
    p, phi0 = small_problem(dims, seed=2)
    theta = adjoint.theta_of(p)
    tt = [theta[0] * 1.1, theta[1] * 0.95, theta[2], theta[3], theta[4] * 1.3]
    out = {}
    # the checkpointed pass with its recomputation on a second stream
    # (overlapping the adjoint steps) and without
    for mode in ("tape", "checkpoints", "checkpoints-overlap"):
        tape = mode == "tape"
        run = adjoint.AdjointRunner(dims, p, checkpoint_every=4, tape=tape,
                                    overlap=mode.endswith("overlap"), **_runner_kw(precision))
        tgt = run.forward(tt, torch.from_numpy(phi0), 23, keep=False)[:7].sum(0).clone()
        L, grads = run.loss_and_grad(theta, torch.from_numpy(phi0), tgt, 23)
        assert (run.tape is not None) == tape
        assert run.tape is not None or (sorted(run.checkpoints) == [0, 4, 8, 12, 16, 20]
                                        and len(run.tail[0]) == 3)
        out[mode] = (L, grads, run.lam0.cpu())
    print(precision, {m: v[1] for m, v in out.items()}, flush=True)
    for mode in ("checkpoints", "checkpoints-overlap"):
        assert out["tape"][0] == out[mode][0]
        assert out["tape"][1] == out[mode][1]
        assert torch.equal(out["tape"][2], out[mode][2])


@pytest.mark.parametrize("precision", ["f64", "f32", "f32/f64", "f32/f32d"])
def test_backward_steps_equal_single_steps(need_gpu, precision):
    """zsc_adj_backward_steps (each step's stencil term left as one field
    for the next step to add) equals one zsc_adj_backward call per step bit
    for bit, in dL/da and in the coefficient sums: a whole 7-step run from
    a_0 (unpulled first state), a 3-step range of pulled states, and a
    single step whose phi the call computes."""
    import torch
    dims = (14, 10, 8)
    p, phi0 = small_problem(dims, seed=7)
    run = adjoint.AdjointRunner(dims, p, checkpoint_every=4, tape=True, **_runner_kw(precision))
    run.forward(adjoint.theta_of(p), torch.from_numpy(phi0), 7)
    states, phis = run.tape
    h = run.ctx._h
    g = torch.Generator(device="cuda").manual_seed(11)
    lam_n = torch.randn((26, run.n), dtype=run.ldtype, device="cuda", generator=g)
    gp = ctypes.cast(run.grads.data_ptr(), ctypes.POINTER(ctypes.c_double))
    cases = [(0, 7, True), (2, 5, True), (0, 1, False)]
    for n0, n1, with_phi in cases:
        out = {}
        for mode in ("single", "steps"):
            run.grads.zero_()
            lam, lam2, lam3 = lam_n.clone(), run._lam(), run._lam()
            torch.cuda.synchronize()
            if mode == "single":
                for n in range(n1 - 1, n0 - 1, -1):
                    phi = _vp(phis[n]) if (with_phi or n > 0) else None
                    check(lib().zsc_adj_backward(h, _vp(states[n]), 0 if n == 0 else 1, phi,
                                                 _vp(lam), _vp(lam2), _vp(run.scratch), gp))
                    lam, lam2 = lam2, lam
                res = lam
            else:
                k = n1 - n0
                sp = (ctypes.c_void_p * k)(*[states[n].data_ptr() for n in range(n0, n1)])
                pp = (ctypes.c_void_p * k)(*[phis[n].data_ptr() if (with_phi or n > 0) else None
                                             for n in range(n0, n1)])
                check(lib().zsc_adj_backward_steps(h, k, sp, pp, 0 if n0 == 0 else 1, _vp(lam),
                                                   _vp(lam2), _vp(lam3), _vp(run.scratch), gp))
                res = lam2
                run.ctx.sync()
                assert torch.equal(lam, lam_n)  # lam_next is not modified
            run.ctx.sync()
            out[mode] = (res.clone(), run.grads.clone())
        d = (out["single"][0] - out["steps"][0]).abs().max().item()
        print(precision, (n0, n1), "max |d lam|", d, flush=True)
        assert torch.equal(out["single"][0], out["steps"][0])
        assert torch.equal(out["single"][1], out["steps"][1])


@pytest.mark.parametrize("overlap", [True, False])
def test_backward_after_a_failed_pass(need_gpu, overlap):
    """A backward pass stopped by a host-side error after its first segment
    leaves that segment's coefficient sums in the scratch buffer; the next
    pass drops them and gives the gradients of a pass that never failed,
    bit for bit."""
    import torch
    dims = (14, 10, 8)
    p, phi0 = small_problem(dims, seed=9)
    theta = adjoint.theta_of(p)
    run = adjoint.AdjointRunner(dims, p, precision="f32", checkpoint_every=4, tape=False,
                                overlap=overlap)
    tgt = run.forward([theta[0] * 1.1] + theta[1:], torch.from_numpy(phi0), 11,
                      keep=False)[:7].sum(0).clone()
    _, ref = run.loss_and_grad(theta, torch.from_numpy(phi0), tgt, 11)
    a = run.forward(theta, torch.from_numpy(phi0), 11)
    _, lam = run.loss(a, tgt)
    steps = run._adjoint_steps

    def failing(*args):
        steps(*args)
        raise RuntimeError("stopped after one segment")

    run._adjoint_steps = failing
    with pytest.raises(RuntimeError):
        run.backward(lam.clone())
    del run._adjoint_steps
    g = run.backward(lam.clone())
    print("after a failed pass", g, "\nreference        ", ref, flush=True)
    assert g == ref


def test_backward_steps_arguments(need_gpu):
    """Bad step counts, missing phi of a pulled state and aliased adjoint
    lattices are refused with ZSC_ERR_ARG before anything runs."""
    import torch
    from lbm_bench_gpu_amd._lib import ZSC_ERR_ARG, ZscError
    dims = (12, 10, 8)
    p, phi0 = small_problem(dims, seed=8)
    run = adjoint.AdjointRunner(dims, p, precision="f32", checkpoint_every=4, tape=True)
    run.forward(adjoint.theta_of(p), torch.from_numpy(phi0), 3)
    states, phis = run.tape
    h = run.ctx._h
    lam, lam2, lam3 = run._lam(), run._lam(), run._lam()
    sp = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in states])
    pp = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in phis])
    nophi = (ctypes.c_void_p * 3)(phis[0].data_ptr(), None, phis[2].data_ptr())
    bad = [(0, sp, pp, 0, lam, lam2, lam3), (3, sp, nophi, 0, lam, lam2, lam3),
           (3, sp, pp, 0, lam, lam, lam3), (3, sp, pp, 0, lam, lam2, None),
           (3, sp, pp, 0, lam, lam2, lam2), (3, sp, pp, 0, lam, lam2, lam)]
    for k, s_, p_, fp, a, b, c in bad:
        with pytest.raises(ZscError) as e:
            check(lib().zsc_adj_backward_steps(h, k, s_, p_, fp, _vp(a), _vp(b),
                                               None if c is None else _vp(c), _vp(run.scratch),
                                               None))
        assert e.value.code == ZSC_ERR_ARG


def _workload_fd(precision, gravity_on):
    import torch
    w = configs.workload(4)
    p = w.params()
    phi0 = torch.from_numpy(zsc.init_phi_host(w.dims, w.bubbles, p.phi_star, p.width))
    theta = adjoint.theta_of(p)
    tt = [theta[0] * 1.2, theta[1] * 0.9, theta[2], theta[3], theta[4] * 1.5]
    fd_run = adjoint.AdjointRunner(w.dims, p, precision="f64", checkpoint_every=20,
                                   gravity_on=gravity_on)
    tgt = fd_run.forward(tt, phi0, w.steps, keep=False)[:7].sum(0).clone()
    run = adjoint.AdjointRunner(w.dims, p, checkpoint_every=20, gravity_on=gravity_on,
                                **_runner_kw(precision))
    L, grads = run.loss_and_grad(theta, phi0, tgt, w.steps)
    fd = []
    for i in range(5):
        h = 1e-4 * abs(theta[i])
        tp = list(theta)
        tm = list(theta)
        tp[i] += h
        tm[i] -= h
        fd.append((fd_run.loss_only(tp, phi0, tgt, w.steps) - fd_run.loss_only(tm, phi0, tgt, w.steps))
                  / (2 * h))
    print(precision, "gravity", gravity_on, "loss", L, "\nadjoint", grads, "\nFD     ", fd, flush=True)
    return grads, fd


SMOOTH_TOL = {"f64": {"kappa": 1e-6, "beta": 1e-6, "tau_f": 1e-6, "tau_g": 1e-6},
              "f32/f64": {"kappa": 1e-4, "beta": 1e-4, "tau_f": 1e-4, "tau_g": 1e-4},
              "f32": {"kappa": 1e-3, "beta": 1e-3, "tau_f": 2e-2, "tau_g": 5e-2},
              "f32/f32d": {"kappa": 1e-6, "beta": 1e-6, "tau_f": 1.5e-4, "tau_g": 1e-4}}


@pytest.mark.parametrize("precision", ["f64", "f32/f64", "f32", "f32/f32d"])
def test_adjoint_workload_smooth_vs_central_differences(need_gpu, precision):
    """The 128^3 workload (8 bubbles R = 12, seed 4, 400 steps, target from
    kappa x1.2, beta x0.9) with gravity switched off, so the loss is smooth
    in every parameter: the adjoint against FP64 central differences (h =
    1e-4 relative).  FP64: 1.8e-8 (kappa), 2.0e-9 (beta), 2.7e-8 (tau_f),
    3.0e-7 (tau_g), bound 1e-6.  FP32 differentiates its own trajectory and
    propagates the adjoint in FP32: 2.0e-4, 3.6e-4, 6.5e-3, 1.8e-2.  Against
    central differences of the FP32 loss itself (h = 1e-2 .. 3e-3, where
    FP32 noise allows) kappa, beta and tau_f agree to 1e-4 .. 2.5e-6, so
    their distance from FP64 is the FP32 trajectory's; tau_g's derivative,
    1.13 against O(1e4 .. 1e5) for the others, sums lambda (g - g_eq), whose
    non-equilibrium part FP32 states (g stored absolutely) carry to ~1e-3
    per cell: ~1.5% off.  FP32 adjoint lattices over FP64 states ("f32/f64":
    FP64 forward and per-cell reverse arithmetic, FP32 lambda storage) meet
    1e-4 for every parameter.  So do FP32 adjoints over f32d states
    ("f32/f32d": every lattice FP32, g stored as deviations from w_i
    rho_mean, the tau_g sums formed without the rest state's part):
    4.4e-7, 1.9e-7, 6.6e-5, 4.6e-5.  The FP32 bounds are those measurements
    with room; the gravity derivative is 0."""
    grads, fd = _workload_fd(precision, False)
    errs = {name: abs(a - r) / abs(r) for name, a, r in list(zip(adjoint.THETA, grads, fd))[:4]}
    print(precision, "relative error vs FD", errs, flush=True)
    assert all(errs[k] <= SMOOTH_TOL[precision][k] for k in errs), errs
    assert grads[4] == 0.0


@pytest.mark.parametrize("precision,tol", [("f64", 5e-4), ("f32/f64", 5e-4), ("f32", 3e-3),
                                           ("f32/f32d", 5e-4)])
def test_adjoint_workload_vs_central_differences(need_gpu, precision, tol):
    """The 128^3 workload as specified (gravity x1.5 in the target).  The
    phi > 0 gravity switch makes the loss piecewise smooth: central
    differences straddle switch flips and scatter by ~3e-4 relative between
    h = 1e-3 and 1e-5, so the bound here is that scatter (FP64) and FP32
    accumulation error on top (FP32).  The exact comparison is against the
    taped FP64 gradient above."""
    grads, fd = _workload_fd(precision, True)
    for name, a, r in zip(adjoint.THETA, grads, fd):
        assert abs(a - r) <= tol * abs(r), (name, a, r)


@pytest.mark.parametrize("precision", ["f64", "f32", "f32/f64", "f32/f32d"])
def test_deferred_coefficient_sums_equal_immediate(need_gpu, precision):
    """Backward calls with grads = NULL keep the coefficient sums in the
    scratch buffer, also when it is fresh (no zsc_adj_seed before); one
    zsc_adj_flush then gives what per-step accumulation gives, and
    zsc_adj_forward refuses to overwrite the kept sums."""
    import torch
    from lbm_bench_gpu_amd._lib import ZSC_ERR_STATE, ZscError
    dims = (14, 10, 8)
    p, phi0 = small_problem(dims, seed=3)
    run = adjoint.AdjointRunner(dims, p, checkpoint_every=4, **_runner_kw(precision))
    h = run.ctx._h
    a = [run.forward(adjoint.theta_of(p), torch.from_numpy(phi0), 2)]
    for _ in range(3):
        nxt = run._lattice()
        torch.cuda.synchronize()
        check(lib().zsc_adj_forward(h, _vp(a[-1]), _vp(nxt), _vp(run.scratch)))
        a.append(nxt)
    run.ctx.sync()
    g = torch.Generator(device="cuda").manual_seed(5)
    lam3 = torch.randn((26, run.n), dtype=run.ldtype, device="cuda", generator=g)
    run.scratch.fill_(0x7f)  # garbage where the deferred sums will live
    gp = ctypes.cast(run.grads.data_ptr(), ctypes.POINTER(ctypes.c_double))
    out = {}
    for mode in ("deferred", "immediate"):
        run.grads.zero_()
        lam, lam2 = lam3.clone(), run._lam()
        torch.cuda.synchronize()  # torch ops above ran on torch's stream
        for n in (2, 1, 0):
            check(lib().zsc_adj_backward(h, _vp(a[n]), 0, None, _vp(lam), _vp(lam2),
                                         _vp(run.scratch), None if mode == "deferred" else gp))
            lam, lam2 = lam2, lam
        if mode == "deferred":
            with pytest.raises(ZscError) as e:
                check(lib().zsc_adj_forward(h, _vp(a[0]), _vp(lam2), _vp(run.scratch)))
            assert e.value.code == ZSC_ERR_STATE
            check(lib().zsc_adj_flush(h, _vp(run.scratch), gp))
        run.ctx.sync()
        out[mode] = (lam.clone(), run.grads.clone())
    assert torch.equal(out["deferred"][0], out["immediate"][0])
    d, i = out["deferred"][1], out["immediate"][1]
    print("deferred", d[:14].tolist(), "\nimmediate", i[:14].tolist(), flush=True)
    # the same FP64 sums in another order
    assert torch.allclose(d, i, rtol=1e-9, atol=1e-12 * float(i.abs().max()))


def test_backward_needs_kept_states(need_gpu):
    """A forward pass with keep=False (as loss_only runs) keeps no states;
    a backward pass after it raises instead of summing over nothing."""
    import torch
    dims = (12, 10, 8)
    p, phi0 = small_problem(dims, seed=6)
    run = adjoint.AdjointRunner(dims, p, precision="f64", checkpoint_every=4)
    theta = adjoint.theta_of(p)
    a = run.forward(theta, torch.from_numpy(phi0), 5)
    L, lam = run.loss(a, torch.zeros(run.n, dtype=torch.float64))
    run.loss_only(theta, torch.from_numpy(phi0), torch.zeros(run.n, dtype=torch.float64), 5)
    with pytest.raises(RuntimeError):
        run.backward(lam)


def test_autograd_function_gives_the_runner_gradients(need_gpu):
    """zsc_loss (the torch.autograd.Function) through torch autograd: the
    loss and theta.grad equal AdjointRunner.loss_and_grad's, and an upstream
    gradient scales them (here 3 L)."""
    import torch
    dims = (12, 10, 8)
    p, phi0 = small_problem(dims, seed=7)
    theta0 = adjoint.theta_of(p)
    tt = [theta0[0] * 1.2, theta0[1] * 0.9, theta0[2], theta0[3], theta0[4] * 1.5]
    run = adjoint.AdjointRunner(dims, p, precision="f64", checkpoint_every=3)
    phi0_t = torch.from_numpy(phi0)
    tgt = run.forward(tt, phi0_t, 7, keep=False)[:7].sum(0).clone()
    L_ref, g_ref = run.loss_and_grad(theta0, phi0_t, tgt, 7)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    L = adjoint.zsc_loss(theta, phi0_t, tgt, run, 7)
    (3.0 * L).backward()
    Lv = float(L.detach())
    print("autograd", Lv, theta.grad.tolist(), "\nrunner  ", L_ref, g_ref, flush=True)
    assert Lv == L_ref
    for a, r in zip(theta.grad.tolist(), g_ref):
        assert a == pytest.approx(3.0 * r, rel=1e-14, abs=0.0)


def test_mixed_adjoint_dot_product_and_refusals(need_gpu):
    """FP32 adjoint lattices over FP64 states: one step's <J v, w> (FP64
    central differences) against <v, J^T w> with w stored in FP32 (relative
    1e-6: the FP32 rounding of the stored adjoint), and the lambda precision
    combinations the library refuses."""
    import torch
    from lbm_bench_gpu_amd._lib import ZSC_ERR_PRECISION, PRECISIONS
    dims = (12, 10, 8)
    p, phi0 = small_problem(dims)
    run = adjoint.AdjointRunner(dims, p, precision="f32", states="f64", checkpoint_every=4)
    a = run.forward(adjoint.theta_of(p), torch.from_numpy(phi0), 3)
    n = run.n
    g = torch.Generator(device="cuda").manual_seed(1)
    v = torch.randn((26, n), dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn((26, n), dtype=torch.float64, device="cuda", generator=g).float()
    eps = 1e-6
    ap, am = a + eps * v, a - eps * v
    op, om = run._lattice(), run._lattice()
    lam = run._lam()
    torch.cuda.synchronize()
    check(lib().zsc_adj_forward(run.ctx._h, _vp(ap), _vp(op), _vp(run.scratch)))
    check(lib().zsc_adj_forward(run.ctx._h, _vp(am), _vp(om), _vp(run.scratch)))
    gp = ctypes.cast(run.grads.data_ptr(), ctypes.POINTER(ctypes.c_double))
    check(lib().zsc_adj_backward(run.ctx._h, _vp(a), 0, None, _vp(w), _vp(lam), _vp(run.scratch),
                                 gp))
    run.ctx.sync()
    jv = (op - om) / (2 * eps)
    lhs = float((jv * w.double()).sum())
    rhs = float((v * lam.double()).sum())
    print("mixed dot product", lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-6 * abs(rhs)
    for ctx_prec, lam_prec in (("f32", "f64"), ("f32", "bf16"), ("f64", "bf16"), ("f32d", "f64"),
                               ("f32d", "bf16")):
        with zsc.Context(*dims, precision=ctx_prec) as ctx:
            assert lib().zsc_adj_set_lambda_precision(ctx._h, PRECISIONS[lam_prec]) == ZSC_ERR_PRECISION


def test_deviation_state_adjoint_step_equals_fp32(need_gpu):
    """One adjoint step over an f32d state (g as deviations from w_i rho_mean)
    against the same step over the FP32 state with the rest added back:
    lam_out agrees to FP32 rounding, and so do the coefficient sums, except
    that the deviation form leaves out of inv_tau_rho_one and k0..k2 the rest
    state's part rho_mean sum_x sum_i w_i lam_i (resp. rho_mean sum over its
    population group), which cancels in the tau_g derivative; both are
    checked here, with that part formed in FP64 from the adjoint input."""
    import torch
    dims = (12, 10, 8)
    p, phi0 = small_problem(dims, seed=8)
    W = torch.tensor([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=torch.float64, device="cuda")
    rest = (W.float() * np.float32(p.rho_mean)).double()   # w_i rho_mean as the library forms it
    runs = {}
    for states in ("f32", "f32d"):
        runs[states] = adjoint.AdjointRunner(dims, p, precision="f32", states=states,
                                             checkpoint_every=4)
    a = runs["f32"].forward(adjoint.theta_of(p), torch.from_numpy(phi0), 3)
    runs["f32d"]._set_theta(adjoint.theta_of(p))
    adev = a.clone()
    adev[7:] = (a[7:].double() - rest[:, None]).float()
    g = torch.Generator(device="cuda").manual_seed(9)
    w = torch.randn((26, runs["f32"].n), dtype=torch.float32, device="cuda", generator=g)
    out = {}
    for states, st in (("f32", a), ("f32d", adev)):
        run = runs[states]
        lam = run._lam()
        run.grads.zero_()
        gp = ctypes.cast(run.grads.data_ptr(), ctypes.POINTER(ctypes.c_double))
        torch.cuda.synchronize()
        check(lib().zsc_adj_backward(run.ctx._h, _vp(st), 0, None, _vp(w), _vp(lam),
                                     _vp(run.scratch), gp))
        run.ctx.sync()
        out[states] = (lam.double(), run.grads.clone())
    (l32, g32), (ld, gd) = out["f32"], out["f32d"]
    e_lam = float((l32 - ld).abs().max() / l32.abs().max())
    wg = w[7:].double().sum(1)                              # sum over cells per g population
    part = {6: p.rho_mean * float((W * wg).sum()), 7: p.rho_mean * float(wg[0]),
            8: p.rho_mean * float(wg[1:7].sum()), 9: p.rho_mean * float(wg[7:].sum())}
    print("lam_out rel diff", e_lam, "\nf32 ", g32[:14].tolist(), "\nf32d", gd[:14].tolist(),
          "\nrest part", part, flush=True)
    assert e_lam <= 1e-4
    for j in range(14):
        want = float(g32[j]) - part.get(j, 0.0)
        scale = max(abs(float(g32[j])), abs(part.get(j, 0.0)), 1e-30)
        assert abs(float(gd[j]) - want) <= 1e-4 * scale, (j, float(gd[j]), want)
