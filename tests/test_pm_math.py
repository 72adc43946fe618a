"""The project-owned FP64 elementary functions (pelemp_amd/csrc/pm_math.h)
against 40-digit mpmath values: every routine within a few ulp, special
values as IEEE-754 prescribes."""
import math

import mpmath
import numpy as np
import pytest

from oracle import oracle_py as O

mpmath.mp.dps = 40
RNG = np.random.default_rng(20260)


def ulps(got, ref):
    ref = np.array([float(r) for r in ref])
    sp = np.spacing(np.abs(ref))
    return float(np.max(np.abs(got - ref) / sp))


@pytest.mark.parametrize("fn,gen,ref,bound", [
    (0, lambda: RNG.uniform(-700, 700, 3000), mpmath.exp, 1.0),
    (0, lambda: RNG.uniform(-1, 1, 3000), mpmath.exp, 1.0),
    (1, lambda: np.exp(RNG.uniform(-700, 700, 3000)), mpmath.log, 1.0),
    (1, lambda: RNG.uniform(0.5, 2.0, 3000), mpmath.log, 1.0),
    (2, lambda: np.concatenate([RNG.uniform(-0.9, 3, 1500), 10 ** RNG.uniform(-20, 0, 1500)]), mpmath.log1p, 2.0),
    (4, lambda: 10 ** RNG.uniform(-300, 300, 3000) * np.sign(RNG.uniform(-1, 1, 3000)),
     lambda v: mpmath.cbrt(v) if v > 0 else -mpmath.cbrt(-v), 1.0),
    (5, lambda: RNG.uniform(-100, 100, 3000), mpmath.sin, 1.0),
    (6, lambda: RNG.uniform(-100, 100, 3000), mpmath.cos, 1.0),
    (8, lambda: RNG.uniform(-1, 1, 3000), mpmath.acos, 3.0),
])
def test_unary_ulp(fn, gen, ref, bound):
    x = gen()
    assert ulps(O.pm_eval(fn, x), [ref(mpmath.mpf(float(v))) for v in x]) <= bound


def test_pow_relative_error():
    # exp(y log x): error grows with |y log x|; bound 4 ulp * (1 + |y ln x|)
    x = RNG.uniform(1e-3, 100, 3000)
    y = RNG.uniform(-3, 3, 3000)
    got = O.pm_eval(3, x, y)
    ref = np.array([float(mpmath.power(mpmath.mpf(float(a)), mpmath.mpf(float(b)))) for a, b in zip(x, y)])
    err = np.abs(got - ref) / np.spacing(np.abs(ref))
    assert np.all(err <= 4.0 * (1.0 + np.abs(y * np.log(x))))


def test_atan2_ulp():
    y = RNG.uniform(-5, 5, 3000)
    x = RNG.uniform(-5, 5, 3000)
    got = O.pm_eval(7, y, x)
    assert ulps(got, [mpmath.atan2(mpmath.mpf(float(a)), mpmath.mpf(float(b))) for a, b in zip(y, x)]) <= 3.0


def test_special_values():
    inf = math.inf
    e = O.pm_eval(0, np.array([0.0, 1.0, -1000.0, 1000.0, 710.0]))
    assert e[0] == 1.0 and e[2] == 0.0 and e[3] == inf and e[4] == inf
    assert abs(e[1] - math.e) <= np.spacing(math.e)
    lg = O.pm_eval(1, np.array([1.0, 0.0, -1.0, inf, 5e-324]))
    assert lg[0] == 0.0 and lg[1] == -inf and np.isnan(lg[2]) and lg[3] == inf
    assert abs(lg[4] - math.log(5e-324)) < 1e-12
    cb = O.pm_eval(4, np.array([0.0, -0.0, 8.0, -27.0, 1e-310]))
    assert cb[0] == 0.0 and math.copysign(1, cb[1]) < 0 and cb[2] == 2.0 and cb[3] == -3.0
    assert abs(cb[4] / float(mpmath.cbrt(mpmath.mpf(1e-310))) - 1) < 1e-15  # subnormal input
    p = O.pm_eval(3, np.array([0.0, 0.0, 5.0, 1.0]), np.array([2.0, -1.0, 0.0, 123.4]))
    assert p[0] == 0.0 and p[1] == inf and p[2] == 1.0 and p[3] == 1.0
    l1 = O.pm_eval(2, np.array([0.0, 1e-300, -1.0]))
    assert l1[0] == 0.0 and l1[1] == 1e-300 and l1[2] == -inf


def test_sincos_is_sin_and_cos():
    # one reduction for both: bit for bit the separate functions, in every
    # quadrant, at large arguments and at the special values
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-10.0, 10.0, 20000), rng.uniform(-1e6, 1e6, 2000),
                        np.array([0.0, -0.0, np.pi / 2, np.pi, 3 * np.pi / 2, 1e-300, np.inf, -np.inf, np.nan])])
    s, c = O.pm_eval(5, x), O.pm_eval(6, x)
    s2, c2 = O.pm_eval(11, x), O.pm_eval(12, x)
    assert np.array_equal(s.view(np.uint64), s2.view(np.uint64))
    assert np.array_equal(c.view(np.uint64), c2.view(np.uint64))
