"""Pins of the FP64 oracle (oracle/zsc_oracle.c) -- CPU only.

The reference ships no tests, golden outputs or fixtures for its kernels, so
each oracle function is pinned by what the mathematics fixes: closed forms
derived here from the D3Q7/D3Q19 weights (not by retyping the oracle's
expressions), exact conservation laws, symmetries, brute-force sums, and a
second transcription that keeps the reference's ghost-plane structure
(tests/ghost_ref.py).
"""
import numpy as np
import pytest

import oracle as O
from tests.ghost_ref import GhostBox

# D3Q19 lattice from first principles: rest, 6 faces, 12 edges, in the
# reference's direction order (src/stream_g.cu:56-76)
C = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
              (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
              (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)])
WQ = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CS2 = 1.0 / 3.0


def phys(**kw):
    return O.Physics(**kw)


def shift(a, c):
    """a(x + c) on a periodic (z, y, x) array."""
    return np.roll(a, shift=(-c[2], -c[1], -c[0]), axis=(0, 1, 2))


def lap_weights(phi):
    """Isotropic Laplacian (2/cs^2) sum_i w_i (phi(x+c_i) - phi(x))."""
    return (2.0 / CS2) * sum(WQ[i] * (shift(phi, C[i]) - phi) for i in range(19))


def grad_weights(phi):
    """Isotropic gradient (1/cs^2) sum_i w_i c_i phi(x+c_i)."""
    return [(1.0 / CS2) * sum(WQ[i] * C[i][a] * shift(phi, C[i]) for i in range(19)) for a in range(3)]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ------------------------------------------------------------ derivation --
@pytest.mark.parametrize("sigma,width,phi_star", [(0.01, 4.0, 1.0), (0.03, 5.0, 2.0), (1e-3, 3.0, 0.5)])
def test_derive_zsc_identities(sigma, width, phi_star):
    c = O.derive(phys(sigma=sigma, width=width, phi_star=phi_star))
    beta, kappa = c.alpha, c.kappa
    # sigma = (4/3) sqrt(2 kappa beta) phi*^3 and W = sqrt(2 kappa / beta) / phi*
    assert np.isclose(4.0 / 3.0 * np.sqrt(2 * kappa * beta) * phi_star ** 3, sigma, rtol=1e-14)
    assert np.isclose(np.sqrt(2 * kappa / beta) / phi_star, width, rtol=1e-14)
    assert c.alpha4 == 4 * c.alpha


def test_derive_weights_and_relaxation():
    tau_g, tau_f = 0.8, 0.7
    c = O.derive(phys(tau_rho=tau_g, tau_phi=tau_f))
    # K_i = w_i / tau; the D3Q19 weights sum to one
    assert np.isclose(c.k0 + 6 * c.k1 + 12 * c.k2, 1 / tau_g, rtol=1e-15)
    assert np.isclose(c.k1 / c.k0, 1 / 6) and np.isclose(c.k2 / c.k1, 0.5)
    # KC_i = (1 - 1/(2 tau)) WC_i with WC = (1, 1/6, 1/12)
    assert np.isclose(c.kc0, 1 - 0.5 / tau_g) and np.isclose(c.kc1, (1 - 0.5 / tau_g) / 6)
    assert np.isclose(c.kc2, (1 - 0.5 / tau_g) / 12)
    assert np.isclose(c.eta, 1 / (tau_f + 0.5)) and np.isclose(c.eta + c.eta2, 1.0)
    assert np.isclose(c.inv_eta2, 0.5 / (1 - c.eta))
    assert np.isclose(c.inv_tau_rho_one, 1 - 1 / tau_g)


def test_derive_gravity_from_eo_and_override():
    c = O.derive(phys(rho_l=1.0, rho_h=1000.0, phi_star=-1.0, gravity=-1.0, eo=10.0, r0=10.0,
                      sigma=0.01))
    # src/paramread.c:161-162 with phi* = (rho_h - rho_l)/2
    assert np.isclose(c.phi_star, 499.5)
    assert np.isclose(c.gravity, 0.25 * 10.0 * 0.01 / (999.0 * 100.0))
    assert np.isclose(c.grav, 2 * 499.5 * c.gravity)
    c2 = O.derive(phys(gravity=1e-6, phi_star=1.0))
    assert c2.gravity == 1e-6 and c2.grav == 2e-6 and c2.rho_mean == 5.5


# ------------------------------------------------------------------- init --
def test_init_phi_profile_special_values():
    phi_star, W, R = 1.0, 4.0, 5.0
    shape = (24, 24, 24)
    b = [[12.0, 12.0, 12.0, R]]   # z is 1-based: centre plane 12 is stored plane 11
    phi = O.init_phi(shape, b, phi_star, W)
    assert phi[11, 12, 12] == pytest.approx(phi_star * np.tanh(2 * R / W), rel=1e-15)
    # r = R exactly (offset 5 along x): tanh(0) = 0 exactly
    assert phi[11, 12, 17] == 0.0
    # r = 3,4,0 -> 5 as well
    assert phi[11, 16, 15] == 0.0
    # r = R + W = 9: still inside the reach, tanh(-2)
    assert phi[11, 12, 21] == pytest.approx(-phi_star * np.tanh(2.0), rel=1e-15)
    # just outside the reach: the bulk value exactly
    assert phi[11, 12, 22] == -phi_star and phi[0, 0, 0] == -phi_star


def test_init_phi_z_is_one_based():
    b = [[4.0, 4.0, 1.0, 2.0]]    # centre on reference plane 1 = stored plane 0
    phi = O.init_phi((8, 8, 8), b, 1.0, 2.0)
    assert phi[0, 4, 4] == pytest.approx(np.tanh(2.0), rel=1e-15)
    assert phi[0, 4, 4] > phi[1, 4, 4]
    # no periodic image: the top plane (reference plane 8) is 7 away
    assert phi[7, 4, 4] == -1.0


def test_init_phi_last_bubble_wins():
    b = [[8.0, 8.0, 8.0, 4.0], [11.0, 8.0, 8.0, 4.0]]
    phi = O.init_phi((16, 16, 16), b, 1.0, 4.0)
    # cell x=9 (stored z = 7): r1 = 1, r2 = 2 -> the second bubble's value
    assert phi[7, 8, 9] == pytest.approx(np.tanh(2 * (4 - 2) / 4.0), rel=1e-15)
    b_rev = b[::-1]
    phi_rev = O.init_phi((16, 16, 16), b_rev, 1.0, 4.0)
    assert phi_rev[7, 8, 9] == pytest.approx(np.tanh(2 * (4 - 1) / 4.0), rel=1e-15)


# ------------------------------------------------------- stencil symbols --
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_laplacian_and_gradient_fourier_symbols(axis):
    n = 16
    m = 3
    k = 2 * np.pi * m / n
    coord = np.arange(n, dtype=float)
    s = [1, 1, 1]
    s[2 - axis] = n
    x = coord.reshape(s)
    phi = np.broadcast_to(0.4 + 0.3 * np.sin(k * x), (n, n, n)).copy()
    c = O.derive(phys())
    # init_f: f_1 = Gamma mu / 2 and mu = 4 beta phi (phi^2 - phi*^2) - kappa lap
    f = np.zeros((7, n, n, n))
    O.lib().zo_init_f(n, n, n, O.ctypes.byref(c), O._p(phi), O._p(f))
    mu = 2 * f[1] / c.gamma
    lap = (c.alpha4 * phi * (phi * phi - c.phi_star_sq) - mu) / c.kappa
    assert np.abs(lap - 2 * (np.cos(k) - 1) * (phi - 0.4)).max() < 1e-11
    # gradient through collision_f at u = 0: f_B1 - f_B2 = t (f_A1 - f_A2) + 2 C_f u'_x,
    # u'_x = F_x / (2 rho), F = mu grad(phi)
    rho = np.full_like(phi, 5.5)
    z = np.zeros_like(phi)
    fa = np.random.default_rng(0).standard_normal((7, n, n, n)) * 1e-3
    fb, _ = O.collide_with(c, phi, rho, z, z, z, fa, np.zeros((19, n, n, n)), grav_on=False)
    cf = c.inv_tau_phi * c.inv_eta2 * phi
    comps = []
    for a, (qp, qm) in enumerate(((1, 2), (3, 4), (5, 6))):
        u_prime = (fb[qp] - fb[qm] - c.inv_tau_phi_one * (fa[qp] - fa[qm])) / (2 * cf)
        comps.append(u_prime * 2 * rho / mu)
    for a in range(3):
        want = np.sin(k) * 0.3 * np.cos(k * x) if a == axis else 0.0
        assert np.abs(comps[a] - want).max() < 1e-9


def test_stencils_match_weight_forms_on_random_field():
    n = (10, 12, 14)
    rng = np.random.default_rng(1)
    phi = 0.2 + 0.5 * rng.random(n)
    c = O.derive(phys())
    f = np.zeros((7,) + n)
    O.lib().zo_init_f(n[2], n[1], n[0], O.ctypes.byref(c), O._p(phi), O._p(f))
    mu = 2 * f[1] / c.gamma
    lap = (c.alpha4 * phi * (phi * phi - c.phi_star_sq) - mu) / c.kappa
    assert rel(lap, lap_weights(phi)) < 1e-11
    # f_0 = phi - 3 Gamma mu  (src/init_f.cu:74)
    assert rel(f[0], phi - 3 * c.gamma * mu) < 1e-14


# -------------------------------------------------------------- collision --
def random_state(shape, seed, c):
    rng = np.random.default_rng(seed)
    phi = rng.uniform(-0.9, 0.9, shape)
    st = O.State(shape, c).init(phi)
    st.f += 1e-3 * rng.standard_normal(st.f.shape)
    st.g += 1e-3 * rng.standard_normal(st.g.shape)
    return st


def test_collision_conserves_phi_and_rho_per_cell():
    c = O.derive(phys())
    st = random_state((6, 7, 8), 2, c)
    fb, gb = O.collide(c, st.f, st.g)
    phi, rho, *_ = st.moments()
    assert np.abs(fb.sum(0) - phi).max() < 1e-14
    assert np.abs(gb.sum(0) - rho).max() < 1e-13


def test_collision_momentum_closed_form():
    """sum_i c_i g_B,i = rho u + F [2/3 + 1/(6 tau)] + (1 - 1/(2 tau)) u' F, component-wise:
    derived from the D3Q19 weights for the reference's forcing, whose axis pairs
    carry u_a (K1 rho + KC1 F_a) (src/collision_g.cu:131,137,143)."""
    c = O.derive(phys(tau_rho=0.9))
    tau = 0.9
    shape = (8, 8, 8)
    st = random_state(shape, 3, c)
    phi, rho, ux, uy, uz = st.moments()
    _, gb = O.collide(c, st.f, st.g, grav_on=True)
    lap = lap_weights(phi)
    mu = c.alpha4 * phi * (phi * phi - c.phi_star_sq) - c.kappa * lap
    grad = grad_weights(phi)
    F = [mu * g for g in grad]
    F[2] = F[2] + np.where(phi > 0, c.grav, 0.0)
    for a, u in enumerate((ux, uy, uz)):
        up = u + F[a] / (2 * rho)
        mom = sum(C[i][a] * gb[i] for i in range(19))
        want = rho * u + F[a] * (2 / 3 + 1 / (6 * tau)) + (1 - 1 / (2 * tau)) * up * F[a]
        assert np.abs(mom - want).max() < 1e-12


def test_equilibrium_is_a_fixed_point():
    c = O.derive(phys())
    shape = (6, 6, 6)
    for v in (-1.0, 1.0):
        phi = np.full(shape, v)
        st = O.State(shape, c).init(phi)
        f0, g0 = st.f.copy(), st.g.copy()
        st.run(5, grav_on=False)
        assert np.abs(st.f - f0).max() < 1e-15 and np.abs(st.g - g0).max() < 1e-14


# -------------------------------------------------------------- streaming --
def test_stream_delta_closed_form():
    c = O.derive(phys())
    shape = (5, 6, 7)
    fb = np.zeros((7,) + shape)
    gb = np.zeros((19,) + shape)
    z, y, x = 2, 3, 4
    fb[:, z, y, x] = 1.0
    gb[:, z, y, x] = 1.0
    fa, ga = O.stream(c, fb, gb)
    assert fa[0, z, y, x] == 1.0
    for i in range(1, 7):
        d = C[i]
        tgt = ((z + d[2]) % 5, (y + d[1]) % 6, (x + d[0]) % 7)
        assert fa[i][tgt] == pytest.approx(c.eta, rel=1e-15)
        assert fa[i, z, y, x] == pytest.approx(c.eta2, rel=1e-15)
        assert np.count_nonzero(fa[i]) == 2
    for i in range(19):
        d = C[i]
        tgt = ((z + d[2]) % 5, (y + d[1]) % 6, (x + d[0]) % 7)
        assert ga[i][tgt] == 1.0 and np.count_nonzero(ga[i]) == 1


def test_stream_wraps_every_face():
    c = O.derive(phys())
    shape = (3, 4, 5)
    rng = np.random.default_rng(4)
    gb = rng.standard_normal((19,) + shape)
    fb = rng.standard_normal((7,) + shape)
    fa, ga = O.stream(c, fb, gb)
    for i in range(19):
        assert np.array_equal(ga[i], shift(gb[i], -C[i]))
    for i in range(1, 7):
        assert np.allclose(fa[i], c.eta * shift(fb[i], -C[i]) + c.eta2 * fb[i], rtol=0, atol=1e-15)


# ------------------------------------------------- whole-step properties --
def test_exact_conservation_over_100_steps():
    c = O.derive(phys())
    shape = (10, 12, 8)
    st = random_state(shape, 5, c)
    phi0, rho0, *_ = st.moments()
    st.run(100)
    phi, rho, *_ = st.moments()
    n = phi.size
    assert abs(phi.sum() - phi0.sum()) < 1e-12 * n
    assert abs(rho.sum() - rho0.sum()) < 1e-12 * n * 5.5


def test_translation_covariance_bitwise():
    c = O.derive(phys())
    shape = (8, 9, 10)
    st = random_state(shape, 6, c)
    f, g = st.f.copy(), st.g.copy()
    sh = (1, 2, 3)
    st2 = O.State(shape, c)
    st2.f = np.ascontiguousarray(np.roll(f, sh, axis=(1, 2, 3)))
    st2.g = np.ascontiguousarray(np.roll(g, sh, axis=(1, 2, 3)))
    st.run(4)
    st2.run(4)
    assert np.array_equal(np.roll(st.f, sh, axis=(1, 2, 3)), st2.f)
    assert np.array_equal(np.roll(st.g, sh, axis=(1, 2, 3)), st2.g)


def test_mirror_symmetry_x():
    """x -> -x maps direction i to its mirror image.  Without interfacial force
    the step commutes with it; with force it does not, because the reference's
    axis pairs add KC1 u_a F_a (even under the mirror) with opposite signs to
    g_a+ and g_a- (src/collision_g.cu:131)."""
    shape = (6, 7, 8)
    mirror = [int(np.where((C == C[i] * np.array([-1, 1, 1])).all(1))[0][0]) for i in range(19)]
    fm = mirror[:7]

    def flip(a, perm):
        return np.ascontiguousarray(np.roll(a[perm][..., ::-1], 1, axis=-1))

    def run_pair(c):
        st = random_state(shape, 7, c)
        st2 = O.State(shape, c)
        st2.f, st2.g = flip(st.f, fm), flip(st.g, mirror)
        st.run(3)
        st2.run(3)
        return rel(flip(st.f, fm), st2.f), rel(flip(st.g, mirror), st2.g)

    ef, eg = run_pair(O.derive(phys(sigma=0.0, gravity=0.0)))
    assert ef < 1e-13 and eg < 1e-13
    ef, eg = run_pair(O.derive(phys(sigma=0.05, gravity=0.0)))
    assert eg > 1e-9


@pytest.mark.parametrize("shape,grav", [((6, 8, 8), True), ((4, 6, 10), False)])
def test_matches_ghost_plane_transcription(shape, grav):
    """Oracle (periodic wrap, pull) vs. tests/ghost_ref.py (ghost planes, nb
    tables, push streaming, pack/unpack self-exchange)."""
    c = O.derive(phys(gravity=1e-3))
    st = random_state(shape, 8, c)
    nz, ny, nx = shape
    gb = GhostBox(nx, ny, nz, c)
    gb.set_real(st.f, st.g)
    for _ in range(5):
        gb.step(grav_on=grav)
    st.run(5, grav_on=grav)
    f2, g2 = gb.real()
    assert rel(f2, st.f) < 1e-14 and rel(g2, st.g) < 1e-14


# ------------------------------------------------------- moments / stats --
def test_moments_brute_force():
    c = O.derive(phys())
    st = random_state((5, 6, 7), 9, c)
    phi, rho, ux, uy, uz = st.moments()
    assert np.allclose(phi, st.f.sum(0), rtol=0, atol=1e-14)
    assert np.allclose(rho, st.g.sum(0), rtol=0, atol=1e-13)
    for a, u in enumerate((ux, uy, uz)):
        assert np.allclose(u, np.tensordot(C[:, a], st.g, axes=(0, 0)) / rho, rtol=0, atol=1e-14)


def test_stats_brute_force():
    c = O.derive(phys())
    shape = (6, 9, 10)
    st = random_state(shape, 10, c)
    phi, rho, ux, uy, uz = st.moments()
    s = O.stats(c, phi, rho, ux, uy, uz)
    inner = np.zeros(shape, bool)
    inner[:, 1:-1, 1:-1] = True
    bub = inner & (phi >= 0)
    assert s[0] == pytest.approx(phi.sum(), rel=1e-13)
    assert s[1] == pytest.approx(rho.sum(), rel=1e-13)
    assert s[2] == bub.sum() and s[9] == (phi >= 0).sum()
    assert s[3] == pytest.approx(ux[bub].sum(), rel=1e-12)
    assert s[6] == pytest.approx((ux ** 2 + uy ** 2 + uz ** 2)[inner].max(), rel=1e-14)
    gx, gy, gz = grad_weights(phi)
    pres = (c.alpha * (phi ** 2 * (3 * phi ** 2 - 2 * c.phi_star_sq) - c.phi_star_sq ** 2)
            - c.kappa * (phi * lap_weights(phi) - 0.5 * (gx ** 2 + gy ** 2 + gz ** 2)) + rho / 3)
    assert s[7] == pytest.approx(pres[bub].sum(), rel=1e-10)
    assert s[8] == pytest.approx(pres[inner & (phi < 0)].sum(), rel=1e-10)


# -------------------------------------------------------------- pressure --
def double_well(c, phi):
    """Bulk free-energy density psi = beta (phi^2 - phi*^2)^2 and its phi-derivative."""
    d = phi ** 2 - c.phi_star_sq
    return c.alpha * d * d, 4.0 * c.alpha * phi * d


@pytest.mark.parametrize("phi0", [1.0, -1.0, 0.3, -0.7, 0.0])
def test_pressure_uniform_is_bulk_equation_of_state(phi0):
    """For uniform phi the stencils vanish and the pressure is the bulk
    equation of state p0 = phi psi'(phi) - psi(phi) + rho c_s^2; at phi = +-phi*
    only the ideal part rho c_s^2 remains."""
    c = O.derive(phys())
    shape = (4, 5, 6)
    phi = np.full(shape, phi0)
    rho = np.random.default_rng(1).uniform(1, 10, shape)
    psi, dpsi = double_well(c, phi)
    p = O.pressure(c, phi, rho)
    assert np.allclose(p, phi * dpsi - psi + rho * CS2, rtol=1e-14, atol=1e-16)
    if abs(phi0) == 1.0:
        assert np.allclose(p, rho * CS2, rtol=1e-15, atol=0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pressure_gradient_terms_on_a_fourier_mode(axis):
    """phi = phi0 + A cos(k x_a): the 19-point Laplacian has symbol 2(cos k - 1)
    and the isotropic gradient i sin k, so the Korteweg part of the pressure
    is -kappa (phi lap phi - |grad phi|^2 / 2) with those closed forms."""
    c = O.derive(phys(sigma=0.05))
    shape = [8, 8, 8]
    n = 16
    shape[2 - axis] = n
    k = 2 * np.pi * 3 / n
    coord = np.indices(shape)[2 - axis]
    phi0, A = 0.2, 0.6
    phi = phi0 + A * np.cos(k * coord)
    rho = np.full(shape, 5.5)
    lap = 2 * (np.cos(k) - 1) * A * np.cos(k * coord)
    grad = -A * np.sin(k) * np.sin(k * coord)
    psi, dpsi = double_well(c, phi)
    expect = phi * dpsi - psi - c.kappa * (phi * lap - 0.5 * grad ** 2) + rho * CS2
    assert np.allclose(O.pressure(c, phi, rho), expect, rtol=1e-13, atol=1e-15)


def test_pressure_sums_of_stats():
    c = O.derive(phys())
    shape = (6, 9, 10)
    st = random_state(shape, 11, c)
    phi, rho, ux, uy, uz = st.moments()
    s = O.stats(c, phi, rho, ux, uy, uz)
    p = O.pressure(c, phi, rho)
    inner = np.zeros(shape, bool)
    inner[:, 1:-1, 1:-1] = True
    assert s[7] == pytest.approx(p[inner & (phi >= 0)].sum(), rel=1e-13)
    assert s[8] == pytest.approx(p[inner & (phi < 0)].sum(), rel=1e-13)


# ------------------------- phi*mu terms and the f-collision amplitudes --
# Each pin below is evaluated once on the oracle as it is and once with a
# planted mistake in the term it pins (oracle.mutation); the mistake must
# make it fail, which shows the check can see that term.

def mu_weights(c, phi):
    """Chemical potential mu = 4 beta phi (phi^2 - phi*^2) - kappa lap(phi)
    with the Laplacian in its weight form."""
    return 4.0 * c.alpha * phi * (phi * phi - c.phi_star_sq) - c.kappa * lap_weights(phi)


def init_moments(c, phi, rho):
    """Moments of init_f / init_g: (sum f, sum c f, sum c c f), same for g."""
    n = phi.shape
    f = np.zeros((7,) + n)
    g = np.zeros((19,) + n)
    O.lib().zo_init_f(n[2], n[1], n[0], O.ctypes.byref(c), O._p(phi), O._p(f))
    O.lib().zo_init_g(n[2], n[1], n[0], O.ctypes.byref(c), O._p(phi), O._p(rho), O._p(g))

    def mom(d, q):
        m0 = d.sum(0)
        m1 = [np.tensordot(C[:q, a], d, axes=(0, 0)) for a in range(3)]
        m2 = [[np.tensordot(C[:q, a] * C[:q, b], d, axes=(0, 0)) for b in range(3)] for a in range(3)]
        return m0, m1, m2

    return mom(f, 7), mom(g, 19)


def init_moment_errors(c, phi, rho):
    """Largest deviation of the initial equilibria from the moments of the
    ZSC equilibria at u = 0 (Zheng, Shu & Chew, J. Comput. Phys. 218 (2006)):
    order parameter  sum f = phi,  sum c f = 0,  sum c_a c_b f = Gamma mu d_ab;
    momentum         sum g = rho,  sum c g = 0,  sum c_a c_b g = (rho/3 + phi mu) d_ab."""
    mu = mu_weights(c, phi)
    (f0, f1, f2), (g0, g1, g2) = init_moments(c, phi, rho)
    ef = [np.abs(f0 - phi).max()] + [np.abs(m).max() for m in f1]
    eg = [np.abs(g0 - rho).max() / rho.max()] + [np.abs(m).max() for m in g1]
    for a in range(3):
        for b in range(3):
            ef.append(np.abs(f2[a][b] - (c.gamma * mu if a == b else 0.0)).max())
            want = rho / 3 + phi * mu if a == b else 0.0
            eg.append(np.abs(g2[a][b] - want).max())
    return max(ef), max(eg)


def test_init_equilibria_moments_with_nonzero_mu():
    """Random phi (mu of order 1e-2) and random rho: every phi*mu term of
    init_g (src/init_g.cu:77-91) and the Gamma mu amplitude of init_f
    (src/init_f.cu:72-78) enters one of these moments."""
    shape = (6, 7, 8)
    rng = np.random.default_rng(21)
    phi = rng.uniform(-1.2, 1.2, shape)
    rho = rng.uniform(1.0, 10.0, shape)
    c = O.derive(phys(gamma=0.7, sigma=0.05))
    assert np.abs(mu_weights(c, phi)).max() > 1e-2
    ef, eg = init_moment_errors(c, phi, rho)
    assert ef < 1e-14 and eg < 1e-14
    for m in (O.MUT_INITG_REST_3PHIMU, O.MUT_INITG_AXIS_PHIMU):
        with O.mutation(m):
            assert init_moment_errors(c, phi, rho)[1] > 1e-4, m
    with O.mutation(O.MUT_INITF_AF_NO_HALF):
        assert init_moment_errors(c, phi, rho)[0] > 1e-4


def uniform_drift(phi0, steps=4, **kw):
    """Largest change of f and g over `steps` steps from the initial
    equilibria of a uniform phi = phi0 without gravity."""
    c = O.derive(phys(gravity=0.0, **kw))
    shape = (4, 5, 6)
    st = O.State(shape, c).init(np.full(shape, phi0))
    f0, g0 = st.f.copy(), st.g.copy()
    st.run(steps, grav_on=False)
    return float(np.abs(st.f - f0).max()), float(np.abs(st.g - g0).max())


@pytest.mark.parametrize("phi0", [0.3, -0.6, 1.4])
def test_uniform_state_with_nonzero_mu_is_a_fixed_point(phi0):
    """A uniform phi away from 0 and +-phi* has mu = 4 beta phi (phi^2 - phi*^2)
    != 0 but no gradient, hence no force and no flow: the state must not
    change.  The collision leaves f_1..6 = Gamma mu / 2 (pinned by the Fourier
    symbol test) unchanged only if A_f = Gamma mu / (2 tau_phi)
    (src/collision_f.cu:105), f_0 only with 6 A_f, and g only if the phi*mu
    terms of collision_g (src/collision_g.cu:113-194) equal init_g's."""
    kw = dict(tau_phi=0.7, tau_rho=0.8, gamma=0.9, sigma=0.05)
    ef, eg = uniform_drift(phi0, **kw)
    assert ef < 1e-15 and eg < 1e-14
    for m in (O.MUT_AF_NO_TAU, O.MUT_F0_ONE_AF):
        with O.mutation(m):
            assert uniform_drift(phi0, **kw)[0] > 1e-5, m
    for m in (O.MUT_CG_REST_3PHIMU, O.MUT_CG_DIAG_6PHIMU):
        with O.mutation(m):
            assert uniform_drift(phi0, **kw)[1] > 1e-5, m


def collision_f_moment_errors(c, st, grav_on=True):
    """One f collision on a random state against the moment closed forms of
    a BGK relaxation towards the ZSC equilibrium with forcing-shifted velocity
    u' = u + F / (2 rho), F = mu grad(phi) (+ gravity on phi > 0 in z):
      sum c_a f_B       = (1 - 1/tau) sum c_a f_A + phi u'_a / (tau (1 - eta))
      sum c_a c_b f_B   = (1 - 1/tau) sum c_a c_b f_A + Gamma mu d_ab / tau
    The 1/(1 - eta) = 2 invEta2 factor compensates the smoothing of the
    streaming (src/paramread.c:147-148); its macroscopic effect is checked in
    test_phi_advection_speed_closed_form."""
    phi, rho, ux, uy, uz = st.moments()
    fb, _ = O.collide(c, st.f, st.g, grav_on=grav_on)
    mu = mu_weights(c, phi)
    grad = grad_weights(phi)
    F = [mu * gr for gr in grad]
    if grav_on:
        F[2] = F[2] + np.where(phi > 0, c.grav, 0.0)
    tau = 1.0 / c.inv_tau_phi
    e1, e2 = [], []
    for a, u in enumerate((ux, uy, uz)):
        up = u + F[a] / (2 * rho)
        m_a = np.tensordot(C[:7, a], st.f, axes=(0, 0))
        m_b = np.tensordot(C[:7, a], fb, axes=(0, 0))
        e1.append(np.abs(m_b - ((1 - 1 / tau) * m_a + phi * up / (tau * (1 - c.eta)))).max())
        for b in range(3):
            w = C[:7, a] * C[:7, b]
            s_a = np.tensordot(w, st.f, axes=(0, 0))
            s_b = np.tensordot(w, fb, axes=(0, 0))
            want = (1 - 1 / tau) * s_a + (c.gamma * mu / tau if a == b else 0.0)
            e2.append(np.abs(s_b - want).max())
    return max(e1), max(e2)


def test_collision_f_moments_pin_af_and_cf():
    """First moment fixes C_f = phi invEta2 / tau_phi, second moment fixes
    A_f = Gamma mu / (2 tau_phi) (src/collision_f.cu:105-114), on a random
    state with flow, interfacial force and gravity."""
    c = O.derive(phys(tau_phi=0.7, gamma=0.8, sigma=0.05, gravity=1e-3))
    st = random_state((6, 7, 8), 22, c)
    st.g += 0.02 * np.random.default_rng(23).standard_normal(st.g.shape)  # flow
    e1, e2 = collision_f_moment_errors(c, st)
    assert e1 < 1e-14 and e2 < 1e-14
    for m in (O.MUT_CF_NO_TAU, O.MUT_CF_HALF):
        with O.mutation(m):
            assert collision_f_moment_errors(c, st)[0] > 1e-5, m
    with O.mutation(O.MUT_AF_NO_TAU):
        assert collision_f_moment_errors(c, st)[1] > 1e-5


def advection_speed(tau_phi, u0=0.01, nx=48, steps=240):
    """Speed at which a small cosine mode of phi travels in a uniform flow u0
    along x, fitted from the phase of its Fourier coefficient."""
    c = O.derive(phys(tau_phi=tau_phi, gravity=0.0))
    shape = (3, 3, nx)
    k = 2 * np.pi / nx
    phi = np.broadcast_to(0.02 * np.cos(k * np.arange(nx)), shape).copy()
    st = O.State(shape, c).init(phi)
    # momentum populations at the D3Q19 equilibrium of rho_mean moving at u0
    cu = 3.0 * C[:, 0] * u0
    geq = WQ * c.rho_mean * (1 + cu + 0.5 * cu * cu - 1.5 * u0 * u0)
    st.g = np.ascontiguousarray(np.broadcast_to(geq[:, None, None, None], (19,) + shape))
    phase = []
    for _ in range(steps):
        st.run(1, grav_on=False)
        phase.append(np.angle(np.fft.rfft(st.f.sum(0).mean(axis=(0, 1)))[1]))
    phase = np.unwrap(phase)[steps // 2:]
    t = np.arange(steps // 2, steps)
    return -np.polyfit(t, phase, 1)[0] / k, c


@pytest.mark.parametrize("tau_phi", [0.7, 1.5])
def test_phi_advection_speed_closed_form(tau_phi):
    """Macroscopic pin of C_f.  The collision relaxes the first moment of f
    to phi u / (1 - eta); the smoothed pull f_A,i(x) = eta f_B,i(x - c_i) +
    (1 - eta) f_B,i(x) (src/stream_f.cu:52-59) moves only the eta part, so
    the flux of phi is eta phi u / (1 - eta) and a small mode of phi travels
    at v = eta / (1 - eta) u0 (5 u0 at tau_phi = 0.7, u0 at 1.5).  The
    fitted speed includes the lattice dispersion of a mode of wavelength 48,
    well under the 3% tolerance."""
    v, c = advection_speed(tau_phi)
    want = c.eta / (1 - c.eta) * 0.01
    assert v == pytest.approx(want, rel=0.03)
    with O.mutation(O.MUT_CF_NO_TAU):
        assert abs(advection_speed(tau_phi)[0] / want - 1) > 0.2


def test_fp32_evaluation_tracks_the_fp64_restatement():
    """liboracle_f32.so is the same restatement evaluated in FP32 (the
    reference's kernels compute in float).  On a perturbed two-bubble box,
    12 steps with gravity: its populations, phi and rho stay within FP32
    rounding of the FP64 restatement (rel. 1e-5; measured 2.3e-6 on f and
    7.3e-7 on g), u (momentum sums of g ~ w_i rho over rho) within 1e-4
    (measured 4.4e-5), and they are not the FP64 values (the build really is
    FP32)."""
    from tests.test_gpu_parity import oracle_coef, small_case
    p, phi0 = small_case((20, 12, 10), seed=7)
    c = oracle_coef(p)
    s64 = O.State(phi0.shape, c).init(phi0).run(12)
    s32 = O.State(phi0.shape, c, dtype=np.float32).init(phi0).run(12)
    assert s32.f.dtype == np.float32 and s32.g.dtype == np.float32
    ef = np.abs(s32.f - s64.f).max() / np.abs(s64.f).max()
    eg = np.abs(s32.g - s64.g).max() / np.abs(s64.g).max()
    assert 1e-9 < ef <= 1e-5 and 1e-9 < eg <= 1e-5, (ef, eg)
    for k, (a, b) in enumerate(zip(O.moments(s32.f, s32.g), s64.moments())):
        assert a.dtype == np.float32
        assert np.abs(a - b).max() <= (1e-5 if k < 2 else 1e-4) * np.abs(b).max(), k
