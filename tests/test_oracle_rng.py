"""Pins of the oracle's random streams and deterministic math (CPU)."""
import json
import os

import numpy as np
import pytest

from oracle.philox import philox4x32_10, u01_pair
from oracle.dmath import det_exp, det_log, det_sincos_2pi, box_muller
from oracle import rng

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# Published Random123 known-answer vectors for Philox4x32-10
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_random123_kat(ctr, key, want):
    got = philox4x32_10(np.array(ctr, np.uint32), np.array(key, np.uint32))
    assert tuple(int(v) for v in got) == want


def test_philox_rocrand_vectors():
    """Vectors written by rocRAND's philox4x32_10_engine (tests/golden/make_philox_kat.py)."""
    with open(os.path.join(GOLD, "philox_kat.json")) as f:
        data = json.load(f)
    for rec in data["vectors"]:
        got = philox4x32_10(np.array(rec["ctr"], np.uint32), np.array(rec["key"], np.uint32))
        assert [int(v) for v in got] == rec["out"]


def test_uniform_range_and_moments():
    u0, u1 = rng.uniforms(3, rng.P_MCC, np.arange(200000), 7, 0)
    u = np.concatenate([u0, u1])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.var() - 1.0 / 12.0) < 2e-3
    # 53-bit resolution: values are multiples of 2^-53
    assert np.all(np.floor(u * 2.0 ** 53) == u * 2.0 ** 53)


def test_normals_moments():
    z0, z1 = rng.normals(1, rng.P_INJ_ION, np.arange(200000), 5, 1)
    z = np.concatenate([z0, z1])
    assert abs(z.mean()) < 5e-3
    assert abs(z.var() - 1.0) < 1e-2
    assert abs(np.mean(z ** 4) - 3.0) < 5e-2
    assert abs(np.corrcoef(z0, z1)[0, 1]) < 1e-2


def test_det_exp_accuracy():
    x = np.concatenate([np.linspace(-745, 709, 400001), np.random.default_rng(0).normal(0, 3, 100000)])
    e = det_exp(x)
    r = np.exp(x)
    ok = r > 1e-300
    rel = np.abs(e[ok] - r[ok]) / r[ok]
    assert rel.max() < 4.5e-16
    assert det_exp(np.array([0.0]))[0] == 1.0
    assert det_exp(np.array([-2000.0]))[0] == 0.0
    assert np.isinf(det_exp(np.array([2000.0]))[0])
    assert np.isnan(det_exp(np.array([np.nan]))[0])


def test_det_log_accuracy():
    x = np.exp(np.linspace(-740, 709, 400001))
    l = det_log(x)
    r = np.log(x)
    assert np.max(np.abs(l - r) / np.maximum(np.abs(r), 1e-300)) < 4.5e-16
    near1 = 1.0 + np.linspace(-0.2, 0.2, 100001)
    assert np.max(np.abs(det_log(near1) - np.log(near1))) < 2e-16
    assert det_log(np.array([1.0]))[0] == 0.0
    assert np.isneginf(det_log(np.array([0.0]))[0])
    assert np.isnan(det_log(np.array([-1.0]))[0])


def test_det_sincos():
    u = np.random.default_rng(2).random(200000)
    s, c = det_sincos_2pi(u)
    # extended-precision (x87 long double) reference
    ang = 2 * np.longdouble("3.14159265358979323846264338327950288") * u.astype(np.longdouble)
    assert np.max(np.abs(s - np.sin(ang).astype(np.float64))) < 3e-16
    assert np.max(np.abs(c - np.cos(ang).astype(np.float64))) < 3e-16
    assert np.max(np.abs(s * s + c * c - 1.0)) < 7e-16
    q = np.array([0.0, 0.25, 0.5, 0.75])
    s, c = det_sincos_2pi(q)
    assert np.allclose(s, [0, 1, 0, -1], atol=1e-16) and np.allclose(c, [1, 0, -1, 0], atol=1e-16)


def test_box_muller_zero():
    a, b = box_muller(np.array([0.0]), np.array([0.0]))
    assert a[0] == 0.0 and b[0] == 0.0


@pytest.mark.parametrize("lam", [0.25, 3.0, 16.0, 17.5, 120.0])
def test_poisson_count_moments(lam):
    v = np.array([rng.poisson_count(11, rng.P_INJ_COUNT, s, lam) for s in range(1500)])
    se = np.sqrt(lam / len(v))
    assert abs(v.mean() - lam) < 5 * se
    assert abs(v.var() / lam - 1.0) < 0.2


def test_poisson_count_pmf_small():
    lam = 2.0
    v = np.array([rng.poisson_count(5, rng.P_INJ_COUNT, s, lam) for s in range(4000)])
    from math import exp, factorial
    for k in range(5):
        p = exp(-lam) * lam ** k / factorial(k)
        f = np.mean(v == k)
        assert abs(f - p) < 5 * np.sqrt(p * (1 - p) / len(v))


def test_poisson_count_edge():
    assert rng.poisson_count(0, rng.P_INJ_COUNT, 1, 0.0) == 0
    assert rng.poisson_count(0, rng.P_INJ_COUNT, 1, -3.0) == 0


def test_stream_keys_independent():
    a = rng.block(0, rng.P_MCC, 5, 6, 0)
    b = rng.block(0, rng.P_SEE, 5, 6, 0)
    c = rng.block(1, rng.P_MCC, 5, 6, 0)
    d = rng.block(0, rng.P_MCC, 5, 7, 0)
    assert len({tuple(a), tuple(b), tuple(c), tuple(d)}) == 4
