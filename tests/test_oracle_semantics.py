"""The oracle's arithmetic kit: PCG4D, hash, exp, sin, saturating casts.

PCG4D and hash are pinned by golden vectors from an independent numpy
restatement of the published algorithm (tests/golden/make_pcg4d.py); exp and
sin (shared, bit-for-bit, by the oracle and the HIP kernels) are pinned
against libm in binary64 with an ulp bound.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_ctypes as oc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pcg4d.json")


def test_pcg4d_matches_golden_vectors():
    doc = json.load(open(GOLD))
    ins = np.array(doc["pcg4d_in"], dtype=np.uint32)
    want = np.array(doc["pcg4d_out"], dtype=np.uint32)
    got = oc.pcg4d(ins)
    assert np.array_equal(got, want)
    # rng_f4 = (u >> 8) * 2^-24 is exact in binary32 and lies in [0, 1)
    f4 = (got >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)
    assert np.array_equal(f4.astype(np.float64), np.array(doc["rng_f4_out"]))
    assert f4.min() >= 0.0 and f4.max() < 1.0


def test_pcg4d_streams_are_distinct_and_uniform():
    # the four lanes of one key, epoch, stream, seed and neighbouring streams
    # must not collide; the top-24-bit floats are uniform
    keys = np.arange(200000, dtype=np.uint32)
    ins = np.stack([keys, np.full_like(keys, 17), np.full_like(keys, 1 << 16), np.full_like(keys, 1)], 1)
    out = oc.pcg4d(ins)
    u = (out >> 8).astype(np.float64) / 16777216.0
    for lane in range(4):
        h, _ = np.histogram(u[:, lane], bins=20, range=(0, 1))
        chi2 = ((h - len(keys) / 20) ** 2 / (len(keys) / 20)).sum()
        assert chi2 < 60.0, (lane, chi2)  # 19 dof, p ~ 1e-6
    assert abs(np.corrcoef(u[:, 0], u[:, 2])[0, 1]) < 0.01


def test_hash_matches_golden_vectors():
    doc = json.load(open(GOLD))
    for a, b in zip(doc["hash_in"], doc["hash_out"]):
        assert oc.lib().oracle_hash(a) == b


def ulp_err(got, want):
    got32 = np.float32(got)
    if want == 0.0:
        return 0.0 if got32 == 0.0 else float("inf")
    spacing = np.spacing(np.float32(abs(want)))
    return abs(float(got32) - want) / float(spacing)


def test_expf_within_two_ulp_of_libm():
    xs = np.concatenate([np.linspace(-103.9, 88.7, 20001), np.random.default_rng(1).uniform(-20, 20, 20000),
                         np.array([0.0, -0.0, 1e-8, -1e-8, 0.5, -0.5, 1.0, -1.0, math.log(2), 87.0, -87.0])])
    worst = 0.0
    for x in xs.astype(np.float32):
        want = math.exp(float(x))
        if want < 1.2e-38:  # subnormal results: absolute error bound
            assert abs(oc.expf(float(x)) - want) <= 2 * 1.4e-45
            continue
        worst = max(worst, ulp_err(oc.expf(float(x)), want))
    assert worst <= 2.0, worst
    assert oc.expf(0.0) == 1.0
    assert math.isinf(oc.expf(89.0)) and oc.expf(-104.0) == 0.0
    assert math.isnan(oc.expf(float("nan")))


def test_expf_monotonic_on_a_dense_grid():
    xs = np.linspace(-10.0, 10.0, 200001, dtype=np.float32)
    ys = np.array([oc.expf(float(x)) for x in xs[::7]], dtype=np.float64)
    assert np.all(np.diff(ys) >= 0)


def test_sinf_within_two_ulp_of_libm():
    xs = np.concatenate([np.linspace(-50.0, 50.0, 40001), np.random.default_rng(2).uniform(0, 16, 20000)])
    worst = 0.0
    for x in xs.astype(np.float32):
        want = math.sin(float(x))
        if abs(want) < 1e-6:
            assert abs(oc.sinf(float(x)) - want) < 1e-12 + 4e-8 * abs(float(x))
            continue
        worst = max(worst, ulp_err(oc.sinf(float(x)), want))
    assert worst <= 2.0, worst
    assert oc.sinf(0.0) == 0.0
