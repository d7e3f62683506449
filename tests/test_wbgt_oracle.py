"""WBGT inputs and oracle on the CPU.

The wet-bulb chain (``solweig_project_amd.met``: Lambert-W, saturation vapour
pressure, dewpoint, Romps iteration; ``calculate_wbgt.py:156-715``) is
pinned by library routines and published relations that are independent of
its formulas:

* ``scipy.special.lambertw(x, -1)`` for the PSEM approximation;
* es(T0) = es0, the boiling point at 101325 Pa, and Bolton (1980)'s
  saturation vapour pressure within 0.15 % over 0-40 C;
* the dewpoint's defining property es(Td) = e;
* Tw = T at saturation;
* the thermodynamic wet-bulb equation of the ASHRAE Handbook
  (Fundamentals 2017, ch. 1 eq. 33, its own heat capacities) solved by
  bisection: within 0.02 K over 0-45 C, 5-95 % RH, 90-105 kPa.

The black-globe oracle (``oracle/wbgt.py``) is pinned by the real root of
its quartic, by Tg = Ta when Tmrt = Ta, by ISO 7726's h_cg = 6.3 v^0.6 / D^0.4
for a 0.15 m globe and by ISO 7243's weights. Mistakes planted in the
oracle (on the CPU) fail these checks.
"""
import numpy as np
import pytest

scipy = pytest.importorskip("scipy")
from scipy import optimize, special  # noqa: E402

from oracle import wbgt as WB  # noqa: E402
from solweig_project_amd import met as M  # noqa: E402


def test_lambertw_lower_branch_matches_scipy():
    x = -np.logspace(-14, np.log10(1 / np.e - 1e-3), 4000)
    got = M.lambertw_lower(x)
    ref = special.lambertw(x, -1).real
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-9
    assert np.isnan(M.lambertw_lower(np.array([0.0, 0.5, -0.4]))).all()


def test_saturation_vapour_pressure_published_values():
    assert M.svp_liquid(273.16) == pytest.approx(611.657, rel=1e-14)
    assert abs(M.svp_liquid(373.15) / 101325.0 - 1) < 2e-3          # boiling point
    tc = np.linspace(0.0, 40.0, 41)
    bolton = 611.2 * np.exp(17.67 * tc / (tc + 243.5))
    assert np.max(np.abs(M.svp_liquid(tc + 273.15) / bolton - 1)) < 1.5e-3


def test_dewpoint_is_where_saturation_meets_the_vapour_pressure():
    T = np.repeat(np.linspace(255.0, 320.0, 14), 12)
    rh = np.tile(np.linspace(3.0, 99.0, 12), 14)
    p = np.full(T.shape, 100000.0)
    e = rh / 100.0 * M.svp_liquid(T)
    q = M.q_from_vapour_pressure(p, e)
    td = M.dewpoint_liquid(p, T, q)
    assert np.all(td <= T)
    assert np.max(np.abs(M.svp_liquid(td) / e - 1)) <= 1e-9


def _ashrae_wet_bulb(tc, rh, p):
    """Thermodynamic wet bulb from the ASHRAE psychrometric equation (kJ/kg, C)."""
    pw = rh / 100.0 * M.svp_liquid(tc + 273.15)
    W = 0.621945 * pw / (p - pw)

    def g(tw):
        ps = M.svp_liquid(tw + 273.15)
        Ws = 0.621945 * ps / (p - ps)
        return ((2501 - 2.326 * tw) * Ws - 1.006 * (tc - tw)) / (2501 + 1.86 * tc - 4.186 * tw) - W

    return optimize.brentq(g, -80.0, tc, xtol=1e-10)


def _grid():
    tc = np.array([0.0, 5, 12, 20, 27, 33, 38, 45])
    rh = np.array([5.0, 15, 30, 50, 70, 95])
    p = np.array([90000.0, 101325.0, 105000.0])
    T, R, P = (a.ravel() for a in np.meshgrid(tc, rh, p, indexing="ij"))
    return T, R, P


def test_wet_bulb_matches_the_ashrae_psychrometric_equation():
    T, R, P = _grid()
    got = M.wet_bulb_from_rh(P, T + 273.15, R)
    assert got.converged and got.iterations <= 6
    ref = np.array([_ashrae_wet_bulb(a, b, c) for a, b, c in zip(T, R, P)])
    assert np.max(np.abs(got.tw_c - ref)) <= 0.02
    assert np.all(got.tw_c <= T)


def test_wet_bulb_at_saturation_is_the_air_temperature():
    T = np.array([275.0, 290.0, 300.0, 315.0])
    got = M.wet_bulb_from_rh(np.full(4, 101325.0), T, np.full(4, 100.0))
    assert np.max(np.abs(got.tw_c - (T - 273.15))) <= 1e-9


def test_wet_bulb_table_iterates_as_a_whole():
    """One convergence test over the table (calculate_wbgt.py:671-678): rows agree with
    their lone evaluation to well below the 0.001 K stopping step."""
    T, R, P = _grid()
    table = M.wet_bulb_from_rh(P, T + 273.15, R)
    for j in (0, 17, 60, len(T) - 1):
        alone = M.wet_bulb_from_rh(P[j:j + 1], T[j:j + 1] + 273.15, R[j:j + 1])
        assert alone.iterations <= table.iterations
        assert abs(alone.tw_c[0] - table.tw_c[j]) <= 1e-5


@pytest.mark.parametrize("mistake", ["constant_latent_heat", "dry_heat_capacity"])
def test_planted_wet_bulb_mistakes_fail_the_ashrae_check(monkeypatch, mistake):
    if mistake == "constant_latent_heat":
        monkeypatch.setattr(M, "latent_heat_liquid", lambda T: M.WB_LV0 + 0.0 * T)
    else:
        monkeypatch.setattr(M, "WB_CPV", M.WB_CPD)
    T, R, P = _grid()
    got = M.wet_bulb_from_rh(P, T + 273.15, R)
    monkeypatch.undo()
    ref = np.array([_ashrae_wet_bulb(a, b, c) for a, b, c in zip(T, R, P)])
    assert np.max(np.abs(got.tw_c - ref)) > 0.02


# ---- black globe / WBGT oracle ----------------------------------------------
def _root(a, tmrt, ta):
    b = (tmrt + 273.15) ** 4 + a * (ta + 273.15)
    r = np.roots([1.0, 0.0, 0.0, a, -b])
    r = r[(np.abs(r.imag) < 1e-6 * np.abs(r)) & (r.real > 0)].real
    assert r.size == 1
    return r[0] - 273.15


CASES = [(va, tm, ta) for va in (0.15, 0.5, 1.0, 3.0, 10.0, 25.0)
         for tm, ta in ((-20.0, -10.0), (0.0, 5.0), (20.0, 20.0), (45.0, 30.0), (60.0, 32.0), (80.0, 45.0))]


def test_globe_coefficient_is_iso7726_for_a_15cm_globe():
    va = np.array([0.15, 1.0, 4.0, 12.0], np.float32)
    hcg = WB.globe_coefficient(va).astype(np.float64) * (0.95 * 5.670374419e-8)
    assert np.allclose(hcg, 6.3 * va.astype(np.float64) ** 0.6 / 0.15 ** 0.4, rtol=1e-4)
    assert np.array_equal(WB.globe_coefficient(va), np.array([M.globe_coefficient(v) for v in va], np.float32))


def test_black_globe_is_the_quartic_root():
    for va, tm, ta in CASES:
        a = WB.globe_coefficient(va)
        got = float(WB.black_globe(a, np.float32(tm), np.float32(ta)))
        assert abs(got - _root(float(a), tm, ta)) <= 0.01, (va, tm, ta, got)


def test_black_globe_edges():
    a = WB.globe_coefficient(2.0)
    for t in (-15.0, 10.0, 33.0):
        assert abs(float(WB.black_globe(a, np.float32(t), np.float32(t))) - t) <= 0.01
    # h_cg of a 30 m/s wind overflows 27 a^4 in float32: -inf, as the reference's float32 tensors give
    assert float(WB.black_globe(WB.globe_coefficient(30.0), np.float32(30), np.float32(25))) == -np.inf
    assert np.isnan(WB.black_globe(a, np.float32(np.nan), np.float32(20)))


def test_wbgt_blend_uses_iso7243_weights():
    a = WB.globe_coefficient(1.5)
    tm = np.array([55.0, 55.0, 55.0, 55.0], np.float32)
    sh = np.array([0.0, 0.0999, 0.1, 1.0], np.float32)
    ta, tw = np.float32(31.0), 24.0
    w07 = np.float32(0.7 * tw)
    tg = float(WB.black_globe(a, tm[0], ta))
    got = WB.wbgt(tm, sh, ta, w07, a)
    no_solar = 0.7 * tw + 0.3 * tg
    solar = 0.7 * tw + 0.2 * tg + 0.1 * 31.0
    assert np.allclose(got, [no_solar, no_solar, solar, solar], atol=2e-5)


def test_planted_globe_mistake_fails_the_root_check(monkeypatch):
    real = np.power

    def wrong_power(x, y):              # Tmrt^3 in place of Tmrt^4
        return real(x, np.float32(3)) if float(y) == 4.0 and np.ndim(x) == 0 and float(x) > 200 else real(x, y)

    monkeypatch.setattr(WB.np, "power", wrong_power)
    bad = float(WB.black_globe(WB.globe_coefficient(1.0), np.float32(45), np.float32(30)))
    monkeypatch.undo()
    assert abs(bad - _root(float(WB.globe_coefficient(1.0)), 45.0, 30.0)) > 0.01
