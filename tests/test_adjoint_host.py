"""Host side of the adjoint (CPU): the chain rule from step-coefficient
derivatives to (kappa, beta, tau_f, tau_g, gravity) against central
differences of the coefficient map, and the taped CPU gradient checker
against finite differences."""
import numpy as np
import pytest

from lbm_bench_gpu_amd import adjoint, configs, zsc


def test_chain_rule_matches_finite_differences_of_the_coefficient_map():
    base = zsc.derive_params(configs.physics(32, 32, 32), 6.0)
    theta = adjoint.theta_of(base)
    rng = np.random.default_rng(0)
    wts = rng.standard_normal(16)
    wts[14:] = 0.0

    def L(th):
        c = adjoint.coefficients(th, base).values
        return sum(wts[j] * c[k] for j, k in enumerate(adjoint.SLOTS[:14]))

    got = adjoint.chain_rule(wts, theta, base.phi_star)
    for i in range(5):
        h = max(1e-6 * abs(theta[i]), 1e-4) if i == 4 else 1e-6 * abs(theta[i])
        tp, tm = list(theta), list(theta)
        tp[i] += h
        tm[i] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        assert got[i] == pytest.approx(fd, rel=1e-6, abs=1e-9), adjoint.THETA[i]


def test_coefficients_reproduce_derived_params():
    base = zsc.derive_params(configs.physics(16, 16, 16), 5.0)
    c = adjoint.coefficients(adjoint.theta_of(base), base).values
    for k, v in base.values.items():
        assert c[k] == pytest.approx(v, rel=1e-15, abs=1e-300), k


def test_taped_gradient_matches_finite_differences():
    import torch

    import oracle as O
    from oracle import taped
    dims = (8, 8, 8)
    p = zsc.derive_params(configs.physics(*dims, gravity=1e-3), 3.0)
    phi0 = torch.from_numpy(O.init_phi((8, 8, 8), [[4.2, 3.9, 4.6, 3.0]], p.phi_star, p.width))
    theta = adjoint.theta_of(p)
    tgt = phi0 * 0.95

    def loss(th, grad=False):
        t = [torch.tensor(v, dtype=torch.float64, requires_grad=grad) for v in th]
        c = taped.coefficients(*t, phi_star=p.phi_star, gamma=p.gamma)
        f, g = taped.init_state(phi0, c, p.rho_mean)
        for _ in range(6):
            f, g = taped.step(f, g, c)
        L = ((f.sum(0) - tgt) ** 2).sum()
        return (L, t) if grad else float(L)

    L, t = loss(theta, True)
    g = torch.autograd.grad(L, t)
    for i in range(5):
        h = 1e-5 * abs(theta[i])
        tp, tm = list(theta), list(theta)
        tp[i] += h
        tm[i] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        assert float(g[i]) == pytest.approx(fd, rel=1e-6), i


@pytest.mark.parametrize("kw", [{"tape": "yes"}, {"checkpoint_every": 0}, {"precision": "bf16"}])
def test_runner_rejects_bad_options(kw):
    base = zsc.derive_params(configs.physics(8, 8, 8), 4.0)
    with pytest.raises(ValueError):
        adjoint.AdjointRunner((8, 8, 8), base, **kw)


def test_tape_bytes():
    """The tape of an N-step run holds B_0 .. B_{N-2} (26 populations) and
    phi of every step: 128^3 cells x 400 steps at FP32 take 90.4 GB."""
    run = adjoint.AdjointRunner.__new__(adjoint.AdjointRunner)
    import torch
    run.n, run.dtype = 128 ** 3, torch.float32
    assert run.tape_bytes(400) == (26 * 399 + 400) * 128 ** 3 * 4
    assert round(run.tape_bytes(400) / 1e9, 1) == 90.4
    assert run.tape_bytes(1) == 128 ** 3 * 4
    run.dtype = torch.float64
    assert run.tape_bytes(2) == 28 * 128 ** 3 * 8
