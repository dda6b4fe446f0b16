"""Moments of spherical Bessel functions, their tables and the Filon rule
(Bolt.jl src/bessel/moments.jl, interpolator.jl, integrator.jl) in the CPU
oracle: every known answer of test/testbessel.jl at its own tolerance, the
moments against mpmath quadrature, the two methods against each other where
both hold, and the Filon rule against quadrature of a quadratic."""
import json
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "bessel_moments.json")))

_TABLES = {}


def table(key):
    key = tuple(key)
    if key not in _TABLES:
        _TABLES[key] = po.MomentTable(*key)
    return _TABLES[key]


def err(v, ref, kind):
    ref = float(ref)
    return abs(v / ref - 1) if kind == "rel" else abs(v - ref)


@pytest.mark.parametrize("e", G["J_moment"], ids=lambda e: f"J-x{e['x']}-a{e['alpha']}")
def test_J_moment_known_answers(e):
    """testbessel.jl:9-50: ∫ t^α J_{5/2}(t) dt (the J moments at half-integer
    order are √(2/π) times the spherical moments of order α + ½)"""
    v = po.J_moment_halfint(e["x"], e["n"], e["alpha"], e["method"])
    assert err(v, e["ref"], "rel") < e["rtol"]


@pytest.mark.parametrize("e", G["sph_moment"], ids=lambda e: f"j{e['nu']}-x{e['x']}-m{e['m']}-{e['method']}")
def test_sph_moment_known_answers(e):
    """testbessel.jl:52-153"""
    v = po.sphj_moment(e["x"], e["nu"], e["m"], e["method"])
    assert err(v, e["ref"], "rel") < e["rtol"]


@pytest.mark.parametrize("e", G["table"], ids=lambda e: f"T{e['line']}-x{e['x']}")
def test_table_known_answers(e):
    """testbessel.jl:155-272: in-range spline values, and the Maclaurin and
    asymptotic branches outside the table's range"""
    v = table((e["nu"], e["order"], e["xmin"], e["xmax"], e["n"]))(e["x"])
    assert len(v) == e["order"]
    for i, r in enumerate(e["refs"]):
        assert err(v[i], r, e["kind"]) < e["tol"], (i, v[i], r)


@pytest.mark.parametrize("e", G["filon"], ids=lambda e: f"F{e['line']}-{e['a']}-{e['b']}")
def test_filon_known_answers(e):
    """testbessel.jl:274-305"""
    v = table(e["table"]).filon(e["f"], e["fp"], e["fpp"], e["k"], e["a"], e["b"])
    assert err(v, e["ref"], e["kind"]) < e["tol"]


def test_filon_loop_known_answer():
    """testbessel.jl:289-291: the piecewise form (the moment at k·a carried over
    from the previous interval) gives the same known answer on (1, 2); here the
    oracle's loop over the nodes t = 1, 2"""
    e = next(e for e in G["filon"] if e["line"] == 274 and e["a"] == 1.0 and e["b"] == 2.0)
    T = table(e["table"])
    v = T.filon_sum([1.0, 2.0], [e["f"], 0.0], [e["fp"], 0.0], [e["fpp"], 0.0], [e["k"]])
    assert err(float(v[0]), e["ref"], e["kind"]) < e["tol"]


def _mp_moment(x, nu, m):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    f = lambda t: t**m * mp.sqrt(mp.pi / (2 * t)) * mp.besselj(nu + mp.mpf(1) / 2, t) if t != 0 else mp.mpf(0)
    pts = [mp.mpf(0)] + [mp.mpf(x) * i / max(1, int(x / 2)) for i in range(1, max(1, int(x / 2)) + 1)]
    return float(mp.quad(f, pts))


@pytest.mark.parametrize("nu", [2, 3, 5])
@pytest.mark.parametrize("m", [0, 1, 2, 3])
def test_moments_against_quadrature(nu, m):
    """I_m(x) = ∫_0^x t^m j_ν(t) dt by 30-digit quadrature: the extended
    series below x = 40 and the asymptotic from 40 on, as the tables (ν ≤ 3)
    use them; the asymptotic's error grows like (ν/x)²ʲ, so at ν = 5 it is
    checked further out"""
    for x in (0.3, 17.0, 26.0, 39.5):
        ref = _mp_moment(x, nu, m)
        assert abs(po.sphj_moment(x, nu, m, "series") - ref) < 1e-15 * max(1.0, abs(ref)), x
    for x in ((40.0, 90.0) if nu <= 3 else (90.0,)):
        ref = _mp_moment(x, nu, m)
        assert abs(po.sphj_moment(x, nu, m, "asymp") - ref) < 2e-15 * max(1.0, abs(ref)), x


def test_series_and_asymptotic_agree_where_both_hold():
    """two independent methods: the extended series and the Lommel asymptotic
    agree between x = 36 and 45 (ν ≤ 3)"""
    for nu in (2, 3):
        for m in range(4):
            for x in np.linspace(36.0, 45.0, 10):
                a, s = po.sphj_moment(x, nu, m, "asymp"), po.sphj_moment(x, nu, m, "series")
                assert abs(a - s) < 3e-15 * max(1.0, abs(s)), (nu, m, x)


def test_maclaurin64_matches_extended_series_at_small_argument():
    for nu in (2, 3):
        for m in range(4):
            for x in (1e-3, 0.05, 0.7, 2.0):
                a, s = po.sphj_moment(x, nu, m, "maclaurin64"), po.sphj_moment(x, nu, m, "series")
                assert abs(a / s - 1) < 1e-14, (nu, m, x)


def test_filon_is_exact_for_a_quadratic():
    """the Filon rule integrates a quadratic against j_ν(kt) exactly up to the
    moments' accuracy: compare with quadrature over intervals inside, below
    and above a table's range"""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    T = table((3, 3, 2.0, 50.0, 4000))
    f = lambda t: 3 * t * t - 0.2 * t + 4
    fp = lambda t: 6 * t - 0.2
    for k, a, b in ((10.0, 0.3, 0.9), (7.0, 1.0, 4.0), (10.0, 0.01, 0.1), (3.0, 30.0, 31.5)):
        g = lambda t: f(t) * mp.sqrt(mp.pi / (2 * k * t)) * mp.besselj(3.5, k * t)
        n = max(2, int((b - a) * k / 2))
        ref = float(mp.quad(g, [a + (b - a) * i / n for i in range(n + 1)]))
        v = T.filon(f(a), fp(a), 6.0, k, a, b)
        assert abs(v - ref) < 1e-9 * max(1.0, abs(ref)), (k, a, b, v, ref)


def test_table_is_exact_at_its_nodes_and_natural_at_its_ends():
    T = table((2, 3, 0.0, 60.0, 601))
    for i in (0, 1, 37, 399, 400, 599, 600):
        x = po.lib().orc_linrange(0.0, 60.0, 601, i)
        v = T(x)
        for m in range(3):
            ref = po.sphj_moment(x, 2, m, "series" if x < 40 else "asymp") if x > 0 else 0.0
            assert abs(v[m] - ref) < 1e-14 * max(1.0, abs(ref)), (i, m)
    c = T.coefficients()
    # Line(OnGrid()) ends: c0 − 2c1 + c2 = 0 at both ends
    for m in range(3):
        assert abs(c[m, 0] - 2 * c[m, 1] + c[m, 2]) < 1e-15 * max(1.0, abs(c[m, 1]))
        assert abs(c[m, -1] - 2 * c[m, -2] + c[m, -3]) < 1e-13 * max(1.0, abs(c[m, -2]))
