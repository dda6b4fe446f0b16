"""Bessel moments, moment tables and the Filon rule on the device (Bolt.jl
src/bessel/moments.jl, interpolator.jl, integrator.jl): every known answer of
test/testbessel.jl at its own tolerance through bolt_sph_moments,
bolt_moment_table_* and bolt_filon; the device tables against the oracle's
coefficient by coefficient; the double-double series against the oracle's
__float128 series; the piecewise Filon sum against the oracle."""
import json
import os

import numpy as np
import pytest

import boltmi
import pyoracle as po
from boltmi import bessel as bs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "bessel_moments.json")))

_TABLES = {}


def dev_table(key):
    key = tuple(key)
    if key not in _TABLES:
        _TABLES[key] = boltmi.sph_bessel_interpolator(*key)
    return _TABLES[key]


def err(v, ref, kind):
    ref = float(ref)
    return abs(v / ref - 1) if kind == "rel" else abs(v - ref)


def test_known_answers_of_the_moment_functions():
    """testbessel.jl:9-153: the asymptotic, the double-double series (which
    stands where the reference sums ₁F₂ by Weniger's transformation) and the
    FP64 Maclaurin series"""
    for e in G["J_moment"]:
        v = bs.J_moment(e["x"], e["n"] + 0.5, e["alpha"], e["method"])
        assert err(v, e["ref"], "rel") < e["rtol"], e
    meth = {"asymp": boltmi.sph_j_moment_asymp, "series": boltmi.sph_j_moment_series,
            "maclaurin64": boltmi.sph_j_moment_maclaurin_1F2}
    for e in G["sph_moment"]:
        v = meth[e["method"]](e["x"], e["nu"], e["m"])
        assert err(v, e["ref"], "rel") < e["rtol"], e


def test_known_answers_of_the_tables():
    """testbessel.jl:155-272, tables of up to 2,000,000 nodes built on the device"""
    for e in G["table"]:
        T = dev_table((e["nu"], e["order"], e["xmin"], e["xmax"], e["n"]))
        v = T(e["x"])
        assert v.shape == (e["order"],) and boltmi.bessel.getorder(T) == e["order"]
        for i, r in enumerate(e["refs"]):
            assert err(v[i], r, e["kind"]) < e["tol"], (e["line"], e["x"], i, v[i], r)


def test_known_answers_of_the_filon_rule():
    """testbessel.jl:274-305, one call per table with all its intervals"""
    by_table = {}
    for e in G["filon"]:
        by_table.setdefault(tuple(e["table"]), []).append(e)
    for key, es in by_table.items():
        T = dev_table(key)
        cols = [np.array([e[c] for e in es]) for c in ("f", "fp", "fpp", "k", "a", "b")]
        v = boltmi.integrate_sph_bessel_filon(*cols, T)
        for e, vi in zip(es, v):
            assert err(vi, e["ref"], e["kind"]) < e["tol"], (e, vi)
        # the scalar form gives the same values
        e0 = es[0]
        assert boltmi.integrate_sph_bessel_filon(e0["f"], e0["fp"], e0["fpp"], e0["k"], e0["a"], e0["b"], T) == v[0]


def test_known_answer_of_the_filon_loop():
    """testbessel.jl:289-291: the piecewise loop over the nodes t = 1, 2
    (bolt_filon_sum) gives the same known answer as the single interval"""
    e = next(e for e in G["filon"] if e["line"] == 274 and e["a"] == 1.0 and e["b"] == 2.0)
    v = boltmi.filon_sum(dev_table(e["table"]), [1.0, 2.0], [e["f"], 0.0], [e["fp"], 0.0], [e["fpp"], 0.0], [e["k"]])
    assert err(float(v[0]), e["ref"], e["kind"]) < e["tol"]


def envelope(x, m):
    """the scale of I_m's terms at x, ~x^{m−1}: FP64 evaluations carry ε of it"""
    return np.maximum(1.0, np.maximum(np.abs(x), 1.0) ** (np.asarray(m) - 1.0))


@pytest.mark.parametrize("key", [(2, 3, 0.0, 300.0, 30001), (3, 4, 0.5, 1000.0, 4001), (1, 3, 5.0, 80.0, 7)])
def test_device_tables_match_the_oracle(key):
    """the device's node values (double-double series below x = 20, exact
    recurrences to 40, asymptotic above) and its chunked prefilter against the
    oracle's __float128 series and sequential prefilter, coefficient by
    coefficient, relative to the moments' envelope ~x^{m−1}"""
    nu, order, lo, hi, n = key
    c_dev = dev_table(key).coefficients()
    c_orc = po.MomentTable(*key).coefficients()
    xn = np.concatenate([[lo], np.linspace(lo, hi, n), [hi]])
    for m in range(order):
        scale = np.maximum(np.abs(c_orc[m]), envelope(xn, m))
        assert np.max(np.abs(c_dev[m] - c_orc[m]) / scale) < 2e-14, (key, m)
    # evaluation inside, below and above the range
    xs = np.concatenate([np.linspace(lo, hi, 97), [lo * 0.5 + 1e-3, hi * 1.7]])
    v_dev = dev_table(key)(xs)
    T = po.MomentTable(*key)
    v_orc = np.array([T(x) for x in xs])
    for m in range(order):
        scale = np.maximum(np.abs(v_orc[:, m]), envelope(xs, m))
        assert np.max(np.abs(v_dev[:, m] - v_orc[:, m]) / scale) < 2e-14, (key, m)


def test_node_method_matches_extended_precision():
    """below x = 20 the double-double series (any ν, real m), from 20 to 40 the
    recurrences down to Si(x), against the oracle's __float128 series"""
    xs = np.concatenate([np.geomspace(1e-3, 1.0, 9), np.linspace(1.5, 19.9, 25)])
    for nu in (1, 2, 3, 6):
        for m in (0, 1, 2, 3, 1.5):
            v = boltmi.sph_j_moment_series(xs, nu, m)
            ref = np.array([po.sphj_moment(x, nu, m, "series") for x in xs])
            assert np.max(np.abs(v / ref - 1)) < 2e-15, (nu, m)
    xs = np.linspace(20.0, 39.99, 120)
    for nu in (1, 2, 3):
        for m in (0, 1, 2, 3):
            v = boltmi.sph_j_moment_series(xs, nu, m)
            ref = np.array([po.sphj_moment(x, nu, m, "series") for x in xs])
            assert np.max(np.abs(v - ref) / np.maximum(np.abs(ref), envelope(xs, m))) < 2e-15, (nu, m)


def test_asymptotic_matches_the_oracle():
    xs = np.linspace(40.0, 5000.0, 41)
    for nu in (1, 2, 3):
        for m in (0, 1, 2, 3):
            v = boltmi.sph_j_moment_asymp(xs, nu, m)
            ref = np.array([po.sphj_moment(x, nu, m, "asymp") for x in xs])
            assert np.max(np.abs(v - ref) / np.maximum(1.0, np.abs(ref))) < 1e-14, (nu, m)


def test_piecewise_filon_sum_matches_the_oracle():
    """∫ f(t) j_3(kt) dt for a smooth f sampled with its derivatives on 400
    intervals, 96 wavenumbers, one wavefront per k on the device"""
    key = (3, 3, 0.0, 2000.0, 200001)
    t = np.linspace(0.05, 12.0, 401)
    f = np.exp(-0.2 * t) * np.cos(0.7 * t) + 0.1 * t
    fp = np.exp(-0.2 * t) * (-0.2 * np.cos(0.7 * t) - 0.7 * np.sin(0.7 * t)) + 0.1
    fpp = np.exp(-0.2 * t) * ((0.04 - 0.49) * np.cos(0.7 * t) + 0.28 * np.sin(0.7 * t))
    ks = np.geomspace(0.01, 150.0, 96)
    v = boltmi.filon_sum(dev_table(key), t, f, fp, fpp, ks)
    ref = po.MomentTable(*key).filon_sum(t, f, fp, fpp, ks)
    assert np.max(np.abs(v - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))


def test_bad_arguments_are_refused():
    with pytest.raises(boltmi._abi.BoltError):
        boltmi.sph_bessel_interpolator(5, 3, 0.0, 10.0, 100)   # ν > 3: no table
    with pytest.raises(boltmi._abi.BoltError):
        boltmi.sph_bessel_interpolator(2, 5, 0.0, 10.0, 100)   # order 3 or 4
    with pytest.raises(boltmi._abi.BoltError):
        boltmi.sph_bessel_interpolator(2, 3, 10.0, 10.0, 100)  # empty range
    T = dev_table((2, 3, 0.0, 50.0, 21))
    with pytest.raises(boltmi._abi.BoltError):
        boltmi.integrate_sph_bessel_filon(1.0, 0.0, 0.0, 0.0, 0.0, 1.0, T)  # k = 0
    with pytest.raises(boltmi._abi.BoltError):
        boltmi.sph_j_moment_asymp(1.0, 3, 0)                  # asymptotic needs x > ν
    with pytest.raises(boltmi._abi.BoltError):
        boltmi.sph_j_moment_series(45.0, 3, 0)                # the node method ends at x = 40
    with pytest.raises(boltmi._abi.BoltError):
        boltmi.sph_j_moment_series(25.0, 3, 1.5)              # recurrences take integer m
