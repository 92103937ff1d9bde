"""Pins the oracle's propagator and force model.

KEP + Yoshida are pinned by the reference's own vectors (SimulationTest 1-s pair,
data/pos_test*.csv) and by closed forms; the other components (no reference test
exercises them) are pinned by the mathematics: numerical gradients of the textbook
potentials, closed forms of the third-body/SRP/drag terms and ephemeris invariants.
"""
import math

import numpy as np
import pytest

import oracle_lib as ol

MU = ol.constant("mu")
RE = ol.constant("r_earth")
L = np.longdouble


def kepler_closed_form(r0, v0, t, mu=MU):
    """Closed-form elliptic Kepler motion: Kepler's equation in the eccentric-anomaly
    difference, then Lagrange f and g coefficients (long double)."""
    r0 = np.array(r0, dtype=L)
    v0 = np.array(v0, dtype=L)
    mu = L(mu)
    rn = np.sqrt(r0 @ r0)
    vn2 = v0 @ v0
    a = 1 / (2 / rn - vn2 / mu)
    n = np.sqrt(mu / a ** 3)
    sigma0 = (r0 @ v0) / np.sqrt(mu)
    # eccentric anomaly form with Lagrange f, g coefficients
    e_cosE0 = 1 - rn / a
    e_sinE0 = sigma0 / np.sqrt(a)
    # solve dE = E - E0: dE - e_cosE0 sin dE + e_sinE0 (1 - cos dE) = n t
    dE = n * L(t)
    for _ in range(100):
        f = dE - e_cosE0 * np.sin(dE) + e_sinE0 * (1 - np.cos(dE)) - n * L(t)
        fp = 1 - e_cosE0 * np.cos(dE) + e_sinE0 * np.sin(dE)
        dE = dE - f / fp
    r = a + (rn - a) * np.cos(dE) + sigma0 * np.sqrt(a) * np.sin(dE)
    F = 1 - a / rn * (1 - np.cos(dE))
    G = a * sigma0 / np.sqrt(mu) * (1 - np.cos(dE)) + rn * np.sqrt(a / mu) * np.sin(dE)
    Fd = -np.sqrt(mu * a) / (r * rn) * np.sin(dE)
    Gd = 1 - a / r * (1 - np.cos(dE))
    return (F * r0 + G * v0).astype(float), (Fd * r0 + Gd * v0).astype(float)


def test_constants_are_the_published_ones():
    # unnormalised J2/C22/S22 from the normalised EGM96 coefficients
    assert math.isclose(ol.constant("j2"), 4.84165371736e-4 * math.sqrt(5), rel_tol=1e-10)
    assert math.isclose(ol.constant("c22"), 2.43914352398e-6 * math.sqrt(5 / 12), rel_tol=1e-8)
    assert math.isclose(ol.constant("s22"), -1.40016683654e-6 * math.sqrt(5 / 12), rel_tol=1e-8)
    assert math.isclose(ol.constant("theta_j2000"), math.radians(280.46061837504), rel_tol=1e-14)
    assert math.isclose(ol.constant("obliquity"), math.radians(84381.448 / 3600.0), rel_tol=1e-14)  # IAU 1976


def test_kepler_1s_reference_vector(kepler_1s):
    """SimulationTest.cpp:39-57: p2 is p1 after one second of Kepler motion."""
    p1, p2 = kepler_1s["p1"], kepler_1s["p2"]
    r, v = ol.propagate([p1["r"]], [p1["v"]], 0.0, kepler_1s["dt"], ol.KEP, 1)
    assert np.max(np.abs(r[0] - p2["r"])) < 1e-9
    assert np.max(np.abs(v[0] - p2["v"])) < 1e-11


def test_pos_test_10s_fixture(pos_test_fixture):
    """data/pos_test.csv + v_test.csv -> *_10s.csv: 10 LEO objects, 10 s of Kepler motion."""
    r0, v0, r1, v1 = pos_test_fixture
    r, v = ol.propagate(r0, v0, 0.0, 10.0, ol.KEP, 1)
    assert np.max(np.abs(r - r1)) < 2e-6
    assert np.max(np.abs(v - v1)) < 2e-9
    assert np.max(np.linalg.norm(r - r1, axis=1) / np.linalg.norm(r1, axis=1)) < 3e-10


def _leo_state():
    a, e = 7000.0, 0.01
    rp = a * (1 - e)
    vp = math.sqrt(MU * (1 + e) / rp)
    inc = math.radians(51.6)
    return np.array([rp, 0.0, 0.0]), np.array([0.0, vp * math.cos(inc), vp * math.sin(inc)])


def test_kepler_matches_closed_form_and_is_fourth_order():
    r0, v0 = _leo_state()
    T = 2000.0
    re, ve = kepler_closed_form(r0, v0, T)
    errs = []
    for dt in (40.0, 20.0, 10.0):
        r, v = ol.propagate([r0], [v0], 0.0, dt, ol.KEP, int(T / dt))
        errs.append(np.linalg.norm(r[0] - re))
    assert errs[-1] < 5e-4
    # 4th order: halving dt divides the error by ~16
    for a, b in zip(errs, errs[1:]):
        assert 11.0 < a / b < 22.0


def test_planted_mistake_breaks_order():
    """A mis-typed Yoshida coefficient must break the 4th-order check above."""
    c, d = ol.yoshida_coefficients()
    assert math.isclose(c.sum(), 1.0, rel_tol=1e-15) and math.isclose(d.sum(), 1.0, rel_tol=1e-15)
    cb = 2 ** (1 / 3)
    assert math.isclose(d[0], 1 / (2 - cb)) and math.isclose(d[1], -cb / (2 - cb))

    def step(r, v, dt, cc, dd):
        for k in range(3):
            r = r + cc[k] * dt * v
            v = v - dd[k] * dt * MU * r / np.linalg.norm(r) ** 3
        return r + cc[3] * dt * v, v

    r0, v0 = _leo_state()
    re, _ = kepler_closed_form(r0, v0, 400.0)
    bad_d = d.copy()
    bad_d[1] *= 1.001
    errs = []
    for dt in (20.0, 10.0):
        r, v = r0.copy(), v0.copy()
        for _ in range(int(400 / dt)):
            r, v = step(r, v, dt, c, bad_d)
        errs.append(np.linalg.norm(r - re))
    assert errs[0] / errs[1] < 8.0


def test_energy_and_momentum_bounded_over_one_day():
    r0, v0 = _leo_state()
    r, v = ol.propagate([r0], [v0], 0.0, 10.0, ol.KEP, 8640)
    e0 = 0.5 * v0 @ v0 - MU / np.linalg.norm(r0)
    e1 = 0.5 * v[0] @ v[0] - MU / np.linalg.norm(r[0])
    assert abs(e1 - e0) / abs(e0) < 1e-9
    h0, h1 = np.cross(r0, v0), np.cross(r[0], v[0])
    assert np.linalg.norm(h1 - h0) / np.linalg.norm(h0) < 1e-12
    re, _ = kepler_closed_form(r0, v0, 86400.0)
    assert np.linalg.norm(r[0] - re) < 1e-2


def test_no_components_is_straight_line():
    r0 = np.array([[7000.0, 1.0, -2.0]])
    v0 = np.array([[0.1, 7.5, 0.3]])
    r, v = ol.propagate(r0, v0, 0.0, 10.0, 0, 25)
    assert np.allclose(r, r0 + 250.0 * v0, rtol=0, atol=1e-9)
    assert np.array_equal(v, v0)


def _num_grad(U, r, h=1e-3):
    g = np.zeros(3)
    for k in range(3):
        e = np.zeros(3, dtype=L)
        e[k] = L(h)
        rl = np.array(r, dtype=L)
        g[k] = float((U(rl + e) - U(rl - e)) / (2 * L(h)))
    return g


def _rotate_to_ef(r, theta):
    c, s = np.cos(L(theta)), np.sin(L(theta))
    return np.array([c * r[0] + s * r[1], -s * r[0] + c * r[1], r[2]], dtype=L)


@pytest.mark.parametrize("r", [[7000.0, 100.0, 50.0], [3000.0, -4000.0, 5500.0], [-100.0, 200.0, 6900.0],
                               [5000.0, 5000.0, 0.0]])
def test_zonal_and_tesseral_terms_are_potential_gradients(r):
    t = 12345.0
    th = ol.lib.or_earth_rotation_angle(t)
    mu, R = L(MU), L(RE)
    j2, c22, s22 = (L(ol.constant(k)) for k in ("j2", "c22", "s22"))

    def U_j2(x):
        r2 = x @ x
        return -j2 * mu * R * R * (3 * x[2] ** 2 - r2) / (2 * r2 ** L(2.5))

    def U_c22(x):
        e = _rotate_to_ef(x, th)
        return 3 * mu * R * R * c22 * (e[0] ** 2 - e[1] ** 2) / (x @ x) ** L(2.5)

    def U_s22(x):
        e = _rotate_to_ef(x, th)
        return 6 * mu * R * R * s22 * e[0] * e[1] / (x @ x) ** L(2.5)

    for bit, U in ((ol.J2, U_j2), (ol.C22, U_c22), (ol.S22, U_s22)):
        a = ol.acceleration(r, [0, 0, 0], t, bit)
        g = _num_grad(U, r)
        assert np.allclose(a, g, rtol=1e-6, atol=1e-6 * np.linalg.norm(g))


def test_j2_closed_form_at_pole_and_equator():
    j2 = ol.constant("j2")
    # equator: a = -3/2 J2 mu R^2 / r^4 radially (outward sign negative = towards Earth)
    a = ol.acceleration([7000.0, 0, 0], [0, 0, 0], 0.0, ol.J2)
    assert math.isclose(a[0], -1.5 * j2 * MU * RE ** 2 / 7000.0 ** 4, rel_tol=1e-13) and a[1] == a[2] == 0
    # pole: a = +3 J2 mu R^2 / r^4 along z (points away from Earth)
    a = ol.acceleration([0, 0, 7000.0], [0, 0, 0], 0.0, ol.J2)
    assert math.isclose(a[2], 3.0 * j2 * MU * RE ** 2 / 7000.0 ** 4, rel_tol=1e-13)


@pytest.mark.parametrize("body,bit,gm", [("sun", ol.SOL, "gm_sun"), ("moon", ol.LUN, "gm_moon")])
def test_third_body_collinear_closed_form(body, bit, gm):
    t = 3.0e7
    s = ol.sun(t) if body == "sun" else ol.moon(t)
    sn = np.linalg.norm(s)
    shat = s / sn
    GM = L(ol.constant(gm))
    for rho in (7000.0, -7000.0):
        r = rho * shat
        a = ol.acceleration(r, [0, 0, 0], t, bit)
        exp = float(GM * (1 / (L(sn) - L(rho)) ** 2 - 1 / L(sn) ** 2) * np.sign(L(sn) - L(rho)))
        assert math.isclose(a @ shat, exp, rel_tol=1e-7)
        assert np.linalg.norm(a - (a @ shat) * shat) < 1e-9 * abs(exp)


def test_srp_inverse_square_and_shadow():
    t = 1.0e8
    s = ol.sun(t)
    shat = s / np.linalg.norm(s)
    aom, cr = 1e-8, 1.3
    r = 7000.0 * shat  # sunlit
    a = ol.acceleration(r, [0, 0, 0], t, ol.SRP, aom=aom, cr=cr)
    d = r - s
    exp = ol.constant("p_sun") * cr * aom * 1e3 * (ol.constant("au") / np.linalg.norm(d)) ** 2
    assert math.isclose(np.linalg.norm(a), exp, rel_tol=1e-12)
    assert np.allclose(a / np.linalg.norm(a), d / np.linalg.norm(d), atol=1e-12)
    # behind the Earth inside the shadow cylinder: no SRP
    assert np.all(ol.acceleration(-7000.0 * shat, [0, 0, 0], t, ol.SRP, aom=aom, cr=cr) == 0)
    # behind the Earth but outside the cylinder: lit
    perp = np.cross(shat, [0, 0, 1.0])
    perp /= np.linalg.norm(perp)
    assert np.linalg.norm(ol.acceleration(-7000.0 * shat + 6500.0 * perp, [0, 0, 0], t, ol.SRP, aom=aom,
                                          cr=cr)) > 0


def test_drag_antiparallel_and_magnitude():
    r = np.array([RE + 400.0, 0.0, 0.0])
    v = np.array([0.0, 7.67, 0.1])
    bc = 0.02
    a = ol.acceleration(r, v, 0.0, ol.DRAG, bc_inv=bc)
    om = ol.constant("omega_earth")
    vrel = v - np.cross([0, 0, om], r)
    rho = 3.725e-12  # Vallado Table 8-4, 400 km row
    exp = 0.5 * bc * rho * (vrel @ vrel) * 1e3
    assert math.isclose(np.linalg.norm(a), exp, rel_tol=1e-12)
    assert np.allclose(a / np.linalg.norm(a), -vrel / np.linalg.norm(vrel), atol=1e-14)


def test_density_table():
    assert ol.lib.or_density(400.0) == 3.725e-12
    assert math.isclose(ol.lib.or_density(425.0), 3.725e-12 * math.exp(-25.0 / 58.515), rel_tol=1e-14)
    assert ol.lib.or_density(1500.0) == pytest.approx(3.019e-15 * math.exp(-500.0 / 268.0), rel=1e-14)
    hs = np.linspace(100, 2000, 500)
    rho = np.array([ol.lib.or_density(h) for h in hs])
    assert np.all(np.diff(rho) < 0)


def test_evasive_preserving_gets_kepler_only():
    r, v = [7000.0, 10.0, 20.0], [0.0, 7.5, 0.1]
    a_full = ol.acceleration(r, v, 0.0, 0xFF, aom=1e-8, bc_inv=0.02, cr=1.2, activity=2)
    a_kep = ol.acceleration(r, v, 0.0, ol.KEP)
    assert np.array_equal(a_full, a_kep)


def test_sun_ephemeris_at_j2000():
    s = ol.sun(0.0)
    ra = math.degrees(math.atan2(s[1], s[0])) % 360
    dec = math.degrees(math.asin(s[2] / np.linalg.norm(s)))
    # almanac: 2000-01-01 12:00 TT, RA 18h45.2m (281.3 deg), Dec -23.03 deg, 0.98333 AU
    assert abs(ra - 281.29) < 0.05
    assert abs(dec + 23.03) < 0.03
    assert abs(np.linalg.norm(s) / ol.constant("au") - 0.98333) < 2e-4


def test_sun_and_moon_invariants():
    for t in np.linspace(0, 3.15e7, 50):
        s, m = ol.sun(t), ol.moon(t)
        au = np.linalg.norm(s) / ol.constant("au")
        assert 0.982 < au < 1.018
        eps = ol.constant("obliquity")
        # Sun stays on the ecliptic: s . n_ecliptic == 0
        n_ecl = np.array([0.0, -math.sin(eps), math.cos(eps)])
        assert abs(s @ n_ecl) / np.linalg.norm(s) < 1e-12
        assert 356000 < np.linalg.norm(m) < 407500
        assert abs(math.degrees(math.asin((m @ n_ecl) / np.linalg.norm(m)))) < 5.35
    # sidereal month: the Moon returns close to its direction after 27.32 days
    m0, m1 = ol.moon(0.0), ol.moon(27.321661 * 86400)
    ang = math.degrees(math.acos(m0 @ m1 / np.linalg.norm(m0) / np.linalg.norm(m1)))
    assert ang < 8.0


# Meeus, Astronomical Algorithms (2nd ed. 1998), Example 47.a: the Moon on 1992 April 12,
# 0h TD (JDE 2448724.5), geometric ecliptic coordinates referred to the mean equinox of date.
MEEUS_47A_JDE = 2448724.5
MEEUS_47A = {"lambda_deg": 133.162655, "beta_deg": -3.229126, "distance_km": 368409.7}


def _ecliptic(m):
    """EME2000 vector -> (longitude, latitude) [deg] on the J2000 ecliptic, distance [km]."""
    eps = ol.constant("obliquity")
    x, y, z = m
    ye = y * math.cos(eps) + z * math.sin(eps)
    ze = -y * math.sin(eps) + z * math.cos(eps)
    r = math.sqrt(x * x + y * y + z * z)
    return math.degrees(math.atan2(ye, x)) % 360.0, math.degrees(math.asin(ze / r)), r


def _meeus_47a_errors(moon_at):
    """Differences of a lunar position function from Meeus Example 47.a.  The series is referred
    to the J2000 equinox (its -1.3972 T term), Meeus to the equinox of date: the longitude is
    carried over by the general precession in longitude p_A = 5029.0966" T + 1.11113" T^2
    (IAU 1976); the ecliptic's own motion (~0.001 deg here) is below the tolerances."""
    t = (MEEUS_47A_JDE - 2451545.0) * 86400.0
    T = t / 3155760000.0
    p_a = (5029.0966 * T + 1.11113 * T * T) / 3600.0
    lam, beta, r = _ecliptic(moon_at(t))
    return (lam - (MEEUS_47A["lambda_deg"] - p_a), beta - MEEUS_47A["beta_deg"], r - MEEUS_47A["distance_km"])


# Tolerances of the low-precision series against the full theory (M&G Sec. 3.3.2 quote a few
# arcminutes and a few hundred km); the oracle lands at 0.005 deg, 0.002 deg and 51 km.
MOON_TOL = {"lambda_deg": 0.1, "beta_deg": 0.03, "distance_km": 150.0}


def _meeus_ok(moon_at):
    dl, db, dr = _meeus_47a_errors(moon_at)
    return abs(dl) < MOON_TOL["lambda_deg"] and abs(db) < MOON_TOL["beta_deg"] and abs(dr) < MOON_TOL["distance_km"]


def _areal_velocity_spread(moon_at, t0=0.0):
    """r^2 dlambda/dt (Kepler's second law) over one sidereal month, relative to its median.
    Solar perturbations make the real Moon's vary by about +-1%."""
    h = []
    for k in range(200):
        t = t0 + k * 27.321661 * 86400.0 / 200
        l0, _, r0 = _ecliptic(moon_at(t - 1800.0))
        l1, _, r1 = _ecliptic(moon_at(t + 1800.0))
        dl = (l1 - l0 + 180.0) % 360.0 - 180.0
        h.append((0.5 * (r0 + r1)) ** 2 * dl)
    h = np.array(h) / np.median(h)
    return h.min(), h.max()


def test_moon_meeus_example_47a():
    dl, db, dr = _meeus_47a_errors(ol.moon)
    assert abs(dl) < MOON_TOL["lambda_deg"], dl
    assert abs(db) < MOON_TOL["beta_deg"], db
    assert abs(dr) < MOON_TOL["distance_km"], dr


def test_moon_areal_velocity_is_nearly_constant():
    for t0 in (0.0, 1.0e8, -3.0e8):
        lo, hi = _areal_velocity_spread(ol.moon, t0)
        assert 0.98 < lo and hi < 1.02, (t0, lo, hi)


@pytest.mark.parametrize("series,term", [("lam", 0), ("lam", 2), ("lam", 3), ("r", 1), ("r", 2),
                                         ("r", 3), ("beta", 0), ("beta", 1)])
def test_planted_lunar_sign_flip_is_caught(series, term):
    """Flipping the sign of a main term of the series (the 22,640", 4,586" evection and 2,370"
    variation terms in longitude, the main distance and latitude terms) must fail the pins."""
    lam, beta, r = ol.moon_series_default()
    amp = {"lam": lam, "beta": beta, "r": r}[series]
    assert amp[term] != 0.0
    amp[term] = -amp[term]
    bad = lambda t: ol.moon_with(t, lam, beta, r)
    caught = not _meeus_ok(bad)
    if series == "lam" and term in (0, 2, 3):
        lo, hi = _areal_velocity_spread(bad)
        caught_invariant = not (0.98 < lo and hi < 1.02)
        assert caught and caught_invariant, (series, term, _meeus_47a_errors(bad), lo, hi)
    assert caught, (series, term, _meeus_47a_errors(bad))
    # the default amplitudes pass the same checks
    assert _meeus_ok(lambda t: ol.moon_with(t, *ol.moon_series_default()))


def test_earth_rotation_sidereal_day():
    th0 = ol.lib.or_earth_rotation_angle(0.0)
    assert math.isclose(th0, math.radians(280.46061837504), rel_tol=1e-14)
    th1 = ol.lib.or_earth_rotation_angle(86164.0905)
    assert abs(th1 - th0) < 1e-5
