"""Pins the oracle's van Genuchten-Mualem closures and face rules.

Golden values: tests/golden/vg_hand_values.json, transcribed from the
reference's tests/unit/test_vangenuchten.cpp:43-126 (1e-12 relative).
Closed forms: FaceConductivity.hpp:19-22 (equal Ks reduces harmonic to the
arithmetic rule; a Ks = 0 flank seals the face).
"""
import json
import math
from pathlib import Path

import pytest

from oracle import oracle

GOLD = json.loads((Path(__file__).parent / "golden" / "vg_hand_values.json").read_text())
SOILS = {k: (v["alpha"], v["n"], v["theta_s"], v["theta_r"], v["aev"])
         for k, v in GOLD["soils"].items()}
KS = {k: v["ks"] for k, v in GOLD["soils"].items()}


def _check(actual, expected, mode):
    if mode == "exact" or expected == 0.0:
        assert actual == expected
    else:
        assert abs(actual / expected - 1.0) <= 1e-12, (actual, expected)


@pytest.mark.parametrize("soil,h,expected,mode", GOLD["theta"])
def test_theta_hand_values(soil, h, expected, mode):
    _check(oracle.vg("theta", SOILS[soil], h), expected, mode)


@pytest.mark.parametrize("soil,h,expected,mode", GOLD["capacity"])
def test_capacity_hand_values(soil, h, expected, mode):
    _check(oracle.vg("capacity", SOILS[soil], h), expected, mode)


@pytest.mark.parametrize("soil,h,expected,mode", GOLD["conductivity"])
def test_conductivity_hand_values(soil, h, expected, mode):
    _check(oracle.vg("conductivity", SOILS[soil], h, KS[soil]), expected, mode)


@pytest.mark.parametrize("soil,wc,expected,mode", GOLD["head"])
def test_head_inversion_hand_values(soil, wc, expected, mode):
    _check(oracle.vg("head", SOILS[soil], wc), expected, mode)


@pytest.mark.parametrize("soil,wc,expected,mode", GOLD["dkdtheta"])
def test_dkdtheta_hand_values(soil, wc, expected, mode):
    _check(oracle.vg("dkdtheta", SOILS[soil], wc, KS[soil]), expected, mode)


def test_head_floor_and_round_trip():
    w = SOILS["warrick"]
    floored = oracle.vg("head", w, 0.0)
    assert floored == oracle.vg("head", w, -1.0)
    assert floored < -1.0e3 and math.isfinite(floored)
    s = SOILS["sand"]
    wc = oracle.vg("theta", s, -2.5)
    assert abs(oracle.vg("head", s, wc) / -2.5 - 1.0) <= 1e-10


def test_dkdtheta_limiter_collapses_near_saturation():
    s = SOILS["sand"]
    a = oracle.vg("dkdtheta", s, 0.459999, KS["sand"])
    assert oracle.vg("dkdtheta", s, 0.4599999, KS["sand"]) == a


def test_capacity_is_derivative_of_theta():
    # C(h) = d theta / dh below air entry (closed form vs central difference)
    for soil in SOILS.values():
        s = (soil[0], soil[1], soil[2], soil[3], -1e-9)
        for h in (-7.0, -1.3, -0.4):
            e = 1e-6
            fd = (oracle.vg("theta", s, h + e) - oracle.vg("theta", s, h - e)) / (2 * e)
            assert abs(oracle.vg("capacity", s, h) - fd) <= 1e-6 * abs(fd) + 1e-12


def test_harmonic_face_rule_closed_forms():
    L = oracle.lib()
    s = SOILS["sand"]
    ks = 1e-5
    for ha, hb in ((-0.3, -0.7), (-2.0, -0.1), (0.1, -1.0)):
        har = L.oracle_face_k_harmonic(*s, ha, ks, 0.1, *s, hb, ks, 0.3)
        ari = 0.5 * (oracle.vg("conductivity", s, ha, ks) + oracle.vg("conductivity", s, hb, ks))
        assert abs(har - ari) <= 1e-14 * ari
    assert L.oracle_face_k_harmonic(*s, -0.3, 0.0, 1.0, *s, -0.3, ks, 1.0) == 0.0
    # saturated series flow: thickness-weighted harmonic mean of Ks
    har = L.oracle_face_k_harmonic(*s, 0.5, 1e-4, 0.2, *s, 0.5, 1e-6, 0.6)
    assert abs(har - 0.8 / (0.2 / 1e-4 + 0.6 / 1e-6)) <= 1e-15
