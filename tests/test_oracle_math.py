"""Pins the oracle's deterministic transcendentals (the float operation
sequences that the GPU library runs too) against float64 numpy: the error in
units of the float32 result's last place, over the argument ranges the
renderer feeds them (angles of a few turns, cosines in [-1, 1], log of
uniforms in (0, 1] and of roughness terms, exp of optical depths), plus the
special values."""
import math

import numpy as np

from oracle.oracle import math_fn


def ulp_err(f, ref):
    sp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
    return np.abs(f.astype(np.float64) - ref) / sp


rng = np.random.default_rng(0)


def test_sin_cos_within_2_ulp():
    x = np.concatenate([np.linspace(-7, 7, 400001), rng.uniform(-100, 100, 100000),
                        np.float32(2 * math.pi) * rng.uniform(0, 1, 100000)]).astype(np.float32)
    for name, fn in (("sin", np.sin), ("cos", np.cos)):
        e = ulp_err(math_fn(name, x), fn(x.astype(np.float64)))
        assert e.max() <= 2.0, (name, e.max())
    assert math_fn("sin", [0.0])[0] == 0 and math_fn("cos", [0.0])[0] == 1


def test_atan2_within_3_5_ulp():
    y = rng.uniform(-1, 1, 400000).astype(np.float32)
    x = rng.uniform(-1, 1, 400000).astype(np.float32)
    x[x == 0] = 0.5
    e = ulp_err(math_fn("atan2", y, x), np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert e.max() <= 3.5, e.max()
    pi = np.float32(math.pi)
    assert math_fn("atan2", [0.0], [1.0])[0] == 0 and math_fn("atan2", [0.0], [-1.0])[0] == pi
    assert math_fn("atan2", [-1e-30], [-1.0])[0] == -pi


def test_acos_within_2_ulp():
    a = np.linspace(-1, 1, 1000001).astype(np.float32)
    r = np.arccos(a.astype(np.float64))
    e = ulp_err(math_fn("acos", a), r)
    assert e.max() <= 2.0, e.max()
    assert math_fn("acos", [1.0])[0] == 0 and math_fn("acos", [-1.0])[0] == np.float32(math.pi)


def test_log_within_1_5_ulp():
    x = np.concatenate([np.exp(rng.uniform(-100, 40, 400000)), rng.uniform(0, 1, 100000),
                        [1e-40, 1e-45, 1.0, 3.4e38]]).astype(np.float32)
    x = x[x > 0]
    e = ulp_err(math_fn("log", x), np.log(x.astype(np.float64)))
    assert e.max() <= 1.5, e.max()
    assert math_fn("log", [1.0])[0] == 0
    out = math_fn("log", [0.0, -1.0])
    assert out[0] == -np.inf and np.isnan(out[1])


def test_exp_within_1_5_ulp():
    x = rng.uniform(-87, 88.7, 400000).astype(np.float32)
    e = ulp_err(math_fn("exp", x), np.exp(x.astype(np.float64)))
    assert e.max() <= 1.5, e.max()
    out = math_fn("exp", [0.0, -87.5, -np.inf, 89.0])
    assert out[0] == 1 and out[1] == 0 and out[2] == 0 and out[3] == np.inf  # no subnormal results


def test_pow_relative_error():
    x = np.exp(rng.uniform(-14, 0, 200000)).astype(np.float32)
    y = rng.uniform(0, 1, 200000).astype(np.float32)
    r = np.power(x.astype(np.float64), y.astype(np.float64))
    assert (np.abs(math_fn("pow", x, y) - r) / r).max() <= 2e-6
