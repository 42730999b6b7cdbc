"""Taped FP64 CPU gradient of the ZSC step -- TEST INFRASTRUCTURE ONLY.

A vectorised restatement of one reference time step (update, collide, stream;
src/main.cu:498-566) in torch float64 on the CPU, differentiable in the five
physical parameters kappa, beta (the reference's alpha), tau_f (tauPhi),
tau_g (tauRho) and gravity.  torch's autograd tape then gives the exact FP64
gradient of a loss of the final state, the yardstick for the hand-written
adjoint of libzsc.  The gravity switch phi > 0 (src/collision_f.cu:96) is a
step function, so its derivative with respect to phi is taken as zero, the
same convention as the adjoint.

Arrays: populations (q, nz, ny, nx); x fastest.  Operand order follows the
reference only loosely: this is a gradient checker, compared at ~1e-10, not a
bitwise oracle.
"""
from __future__ import annotations

import torch

C = torch.tensor([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                  (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
                  (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)])
W = [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12
WC = [1.0] + [1 / 6] * 6 + [1 / 12] * 12


def shift(a, c):
    """a(x + c), periodic."""
    return torch.roll(a, shifts=(-int(c[2]), -int(c[1]), -int(c[0])), dims=(-3, -2, -1))


def coefficients(kappa, beta, tau_f, tau_g, gravity, phi_star, gamma):
    """src/paramread.c:141-170 with kappa, beta, tau_f, tau_g and gravity free."""
    eta = 1.0 / (tau_f + 0.5)
    return dict(kappa=kappa, alpha4=4.0 * beta, phi_star_sq=phi_star * phi_star, gamma=gamma,
                eta=eta, eta2=1.0 - eta, inv_eta2=0.5 / (1.0 - eta), inv_tau_phi=1.0 / tau_f,
                inv_tau_phi_one=1.0 - 1.0 / tau_f, inv_tau_rho_one=1.0 - 1.0 / tau_g,
                inv_tau_g=1.0 / tau_g, kc=1.0 - 0.5 / tau_g, grav=2.0 * phi_star * gravity)


def step(f, g, c, grav_on=True):
    phi = f.sum(0)
    rho = g.sum(0)
    j = [sum(int(C[i][a]) * g[i] for i in range(19) if int(C[i][a]) != 0) for a in range(3)]
    u = [ja / rho for ja in j]
    p = [shift(phi, C[i]) for i in range(19)]
    lap = (2 * sum(p[1:7]) + sum(p[7:19]) - 24.0 * phi) / 6.0
    grad = [sum(int(C[i][a]) * (2.0 if i < 7 else 1.0) * p[i] for i in range(1, 19)
                if int(C[i][a]) != 0) / 12.0 for a in range(3)]
    mu = c["alpha4"] * phi * (phi * phi - c["phi_star_sq"]) - c["kappa"] * lap
    F = [mu * grad[0], mu * grad[1], mu * grad[2]]
    if grav_on:
        F[2] = torch.where(phi > 0, F[2] + c["grav"], F[2])
    v = [u[a] + 0.5 * F[a] / rho for a in range(3)]
    # D3Q7 collision (src/collision_f.cu:104-114)
    af = 0.5 * c["gamma"] * c["inv_tau_phi"] * mu
    cf = c["inv_tau_phi"] * c["inv_eta2"] * phi
    t = c["inv_tau_phi_one"]
    fb = [t * f[0] - 6.0 * af + c["inv_tau_phi"] * phi]
    for i in range(1, 7):
        a = (i - 1) // 2
        s = 1.0 if i % 2 == 1 else -1.0
        fb.append(t * f[i] + af + s * cf * v[a])
    # D3Q19 collision with the reference's forcing (src/collision_g.cu:105-194)
    ut = [3.0 * v[a] for a in range(3)]
    uf = v[0] * F[0] + v[1] * F[1] + v[2] * F[2]
    usq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    tg = c["inv_tau_rho_one"]
    K = [w * c["inv_tau_g"] for w in (W[0], W[1], W[7])]
    KC = [w * c["kc"] for w in (WC[0], WC[1], WC[7])]
    gb = [None] * 19
    gb[0] = tg * g[0] + K[0] * (rho - 6.0 * phi * mu - rho * 9.0 * usq / 6.0) - KC[0] * uf
    t1 = K[1] * (3.0 * phi * mu + rho * (1.0 - 1.5 * usq))
    for a in range(3):
        e1 = t1 + 0.5 * K[1] * rho * ut[a] * ut[a] + KC[1] * (ut[a] * F[a] - uf)
        e2 = ut[a] * (K[1] * rho + KC[1] * F[a])
        gb[1 + 2 * a] = tg * g[1 + 2 * a] + e1 + e2
        gb[2 + 2 * a] = tg * g[2 + 2 * a] + e1 - e2
    t2 = K[2] * (3.0 * phi * mu + rho * (1.0 - 9.0 * usq / 6.0))
    for (ip, im, a, b, sg) in ((7, 8, 0, 1, 1.0), (9, 10, 0, 1, -1.0), (11, 12, 0, 2, 1.0),
                               (13, 14, 0, 2, -1.0), (15, 16, 1, 2, 1.0), (17, 18, 1, 2, -1.0)):
        s = ut[a] + sg * ut[b]
        S = F[a] + sg * F[b]
        e1 = t2 + 0.5 * K[2] * rho * s * s + KC[2] * (s * S - uf)
        e2 = K[2] * rho * s + KC[2] * S
        gb[ip] = tg * g[ip] + e1 + e2
        gb[im] = tg * g[im] + e1 - e2
    # streaming B -> A (src/stream_f.cu:52-59, src/stream_g.cu:56-76 as pulls)
    fa = [fb[0]] + [c["eta"] * shift(fb[i], -C[i]) + c["eta2"] * fb[i] for i in range(1, 7)]
    ga = [shift(gb[i], -C[i]) for i in range(19)]
    return torch.stack(fa), torch.stack(ga)


def init_state(phi, c, rho_mean):
    """init_f / init_g equilibria at u = 0 (src/init_f.cu:69-78, src/init_g.cu:74-91)."""
    p = [shift(phi, C[i]) for i in range(19)]
    lap = (2 * sum(p[1:7]) + sum(p[7:19]) - 24.0 * phi) / 6.0
    mu = c["alpha4"] * phi * (phi * phi - c["phi_star_sq"]) - c["kappa"] * lap
    af = 0.5 * c["gamma"] * mu
    f = torch.stack([phi - 3.0 * c["gamma"] * mu] + [af] * 6)
    rho = torch.full_like(phi, rho_mean)
    ag1 = W[1] * (3.0 * phi * mu + rho)
    ag2 = W[7] * (3.0 * phi * mu + rho)
    g = torch.stack([W[0] * (rho - 6.0 * phi * mu)] + [ag1] * 6 + [ag2] * 12)
    return f, g


def loss_and_grads(f0, g0, theta, fixed, nsteps, phi_tgt, grav_on=True, checkpoint_every=10):
    """L = sum (phi_N - phi_tgt)^2 and dL/d(kappa, beta, tau_f, tau_g, gravity)
    from the autograd tape (segments re-run under torch.utils.checkpoint to
    bound memory)."""
    from torch.utils.checkpoint import checkpoint
    th = [torch.tensor(float(v), dtype=torch.float64, requires_grad=True) for v in theta]
    f, g = f0.clone(), g0.clone()

    def seg(f, g, *th, n):
        c = coefficients(*th, **fixed)
        for _ in range(n):
            f, g = step(f, g, c, grav_on)
        return f, g

    done = 0
    while done < nsteps:
        n = min(checkpoint_every, nsteps - done)
        f, g = checkpoint(lambda a, b, *t, n=n: seg(a, b, *t, n=n), f, g, *th, use_reentrant=False)
        done += n
    loss = ((f.sum(0) - phi_tgt) ** 2).sum()
    grads = torch.autograd.grad(loss, th)
    return float(loss), [float(x) for x in grads]
