"""Mathematical properties that pin the oracle's face physics independently
of the reference dumps: consistency with the exact Euler flux, conservation
under swapped sides, van Leer's supersonic upwinding, the farfield state at
free stream, and the conserved/primitive round trip."""
import numpy as np
import pytest

GAMMA = 1.4


def euler_flux(q, nx, ny, area):
    """Exact normal Euler flux, conserved order [E, rho, rho u, rho v]."""
    rho, u, v, p = q
    un = u * nx + v * ny
    e = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)
    return area * np.array([(e + p) * un, rho * un, rho * u * un + p * nx, rho * v * un + p * ny])


def random_states(n, seed, mach_max=2.5):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        rho = rng.uniform(0.3, 3.0)
        p = rng.uniform(0.2, 4.0)
        a = np.sqrt(GAMMA * p / rho)
        m = rng.uniform(0.0, mach_max)
        th = rng.uniform(0, 2 * np.pi)
        yield np.array([rho, m * a * np.cos(th), m * a * np.sin(th), p])


def unit(seed):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0, 2 * np.pi)
    return np.cos(t), np.sin(t)


@pytest.mark.parametrize("kind", ["roe", "rusanov", "vanleer"])
def test_consistency(kind):
    from oracle.oracle import flux
    for i, q in enumerate(random_states(50, 1)):
        nx, ny = unit(100 + i)
        area = 0.37
        f, ok, _, _ = flux(kind, 0.5, q, q, nx, ny, area)
        assert ok
        assert np.allclose(f, euler_flux(q, nx, ny, area), rtol=1e-12, atol=1e-12)


def test_ausm_consistency_without_face_length():
    """Equal states: AUSM+-up returns the exact flux times |n| (about 1), not
    times the face length (CCFVM/ausm_up_flux.F90:38)."""
    from oracle.oracle import flux
    for i, q in enumerate(random_states(50, 2)):
        nx, ny = unit(200 + i)
        f, _, _, _ = flux("ausm", 0.5, q, q, nx, ny, 0.37)
        assert np.allclose(f, euler_flux(q, nx, ny, 1.0), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("kind", ["roe", "rusanov"])
def test_conservation(kind):
    """F(qL, qR, n) = -F(qR, qL, -n)."""
    from oracle.oracle import flux
    states = list(random_states(40, 3))
    for i in range(0, 40, 2):
        a, b = states[i], states[i + 1]
        nx, ny = unit(300 + i)
        f1, _, _, _ = flux(kind, 0.5, a, b, nx, ny, 0.8)
        f2, _, _, _ = flux(kind, 0.5, b, a, -nx, -ny, 0.8)
        assert np.allclose(f1, -f2, rtol=1e-12, atol=1e-12)


def test_vanleer_supersonic_upwinding():
    """|M| >= 1 on both sides in the same direction: the split flux is the
    exact flux of the upwind state."""
    from oracle.oracle import flux
    rng = np.random.default_rng(4)
    for i in range(30):
        nx, ny = unit(400 + i)
        rho, p = rng.uniform(0.5, 2.0, 2)
        a = np.sqrt(GAMMA * p / rho)
        m = rng.uniform(1.2, 3.0)
        ql = np.array([rho, m * a * nx, m * a * ny, p])
        qr = np.array([rho * 1.3, 1.1 * m * a * nx, 1.1 * m * a * ny, p * 0.9])
        f, _, _, _ = flux("vanleer", 0.5, ql, qr, nx, ny, 0.5)
        assert np.allclose(f, euler_flux(ql, nx, ny, 0.5), rtol=1e-12, atol=1e-12)


def _roe_waves(ql, qr, nx, ny, kq2, shear_sign):
    """Roe-Harten flux by wave decomposition (Roe-averaged state, strengths of
    the acoustic, entropy and shear waves, Harten's fix with delta = 0.1 a).
    kq2 = 0.5, shear_sign = -1 is the textbook form; the reference uses
    kq2 = 1 in the pressure-wave strength (CCFVM/roe_flux.F90:127) and the
    opposite sign of the shear wave's energy term (:91,142)."""
    rl, ul, vl, pl = ql
    rr, ur, vr, pr = qr
    hl = GAMMA * pl / rl / (GAMMA - 1) + 0.5 * (ul ** 2 + vl ** 2)
    hr = GAMMA * pr / rr / (GAMMA - 1) + 0.5 * (ur ** 2 + vr ** 2)
    w = np.sqrt(rr / rl)
    u, v, h = (ul + w * ur) / (1 + w), (vl + w * vr) / (1 + w), (hl + w * hr) / (1 + w)
    q2 = u * u + v * v
    a = np.sqrt((GAMMA - 1) * (h - 0.5 * q2))
    un, ut = u * nx + v * ny, u * ny - v * nx
    d = [rr - rl, rr * ur - rl * ul, rr * vr - rl * vl, (rr * hr - pr) - (rl * hl - pl)]
    dun = (nx * d[1] + ny * d[2] - un * d[0]) / a              # rho du_n / a
    dp = (GAMMA - 1) * (d[3] - u * d[1] - v * d[2] + kq2 * q2 * d[0]) / a ** 2
    s_minus, s_plus = 0.5 * (dp - dun), 0.5 * (dp + dun)
    s_shear = (ny * d[1] - nx * d[2] - ut * d[0]) / a
    s_ent = d[0] - s_minus - s_plus
    lam = np.abs([un - a, un, un, un + a])
    dl = 0.1 * a
    lam = np.where(lam < dl, 0.5 * (dl + lam * lam / dl), lam)
    r_minus = np.array([h - un * a, 1, u - a * nx, v - a * ny])
    r_plus = np.array([h + un * a, 1, u + a * nx, v + a * ny])
    r_ent = np.array([0.5 * q2, 1, u, v])
    r_shear = a * np.array([-shear_sign * ut, 0, ny, -nx])
    diss = (lam[0] * s_minus * r_minus + lam[1] * s_shear * r_shear + lam[2] * s_ent * r_ent
            + lam[3] * s_plus * r_plus)
    return 0.5 * (euler_flux(ql, nx, ny, 1.0) + euler_flux(qr, nx, ny, 1.0) - diss)


def test_roe_is_roe_harten_with_the_reference_deviations():
    from oracle.oracle import flux
    rng = np.random.default_rng(5)
    dev_text = 0.0
    for i in range(40):
        ql = np.array([rng.uniform(.3, 3), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(.2, 4)])
        qr = np.array([rng.uniform(.3, 3), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(.2, 4)])
        nx, ny = unit(700 + i)
        f, ok, _, _ = flux("roe", 0.5, ql, qr, nx, ny, 1.0)
        assert ok
        assert np.allclose(f, _roe_waves(ql, qr, nx, ny, 1.0, 1.0), rtol=1e-12, atol=1e-12)
        dev_text = max(dev_text, np.max(np.abs(f - _roe_waves(ql, qr, nx, ny, 0.5, -1.0))))
    # the textbook form differs at O(1) for general jumps
    assert dev_text > 1e-3


@pytest.mark.parametrize("mach,aoa", [(0.5, 0.0), (0.75, 2.0), (0.3, -2.0)])
def test_farfield_of_free_stream_is_free_stream(mach, aoa):
    from oracle.oracle import farfield_state
    p_inf = 1.0 / (GAMMA * mach * mach)
    a = np.deg2rad(aoa)
    q = np.array([1.0, np.cos(a), np.sin(a), p_inf])
    for i in range(20):
        nx, ny = unit(600 + i)
        qb = farfield_state(mach, aoa, q, nx, ny)
        assert np.allclose(qb, q, rtol=1e-12, atol=1e-12)


def test_rusanov_converts_its_arguments():
    """The reference's Rusanov routine leaves its state arguments in conserved
    variables (CCFVM/rusanov_flux.F90:205-208), which the farfield ghost then
    inherits."""
    from oracle.oracle import flux
    q = np.array([1.2, 0.3, -0.1, 2.0])
    _, _, ql, qr = flux("rusanov", 0.5, q, q, 1.0, 0.0, 1.0)
    e = q[3] / (GAMMA - 1) + 0.5 * q[0] * (q[1] ** 2 + q[2] ** 2)
    assert np.allclose(ql, [e, q[0], q[0] * q[1], q[0] * q[2]], rtol=1e-14)
