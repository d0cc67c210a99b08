"""The skin tuner's kinematic bound (mfixdem.autotune.safe_interval): the
interval it allows is the longest for which two of the smallest spheres
approaching head-on at the largest speed cannot close the skin, checked by
stepping that worst case directly."""
import pytest

import _paths  # noqa: F401
from mfixdem import autotune


@pytest.mark.parametrize("f,r,v,dt", [(1.2, 0.25e-3, 1.0, 1e-5), (1.5, 0.5e-3, 0.3, 2e-5),
                                      (1.1, 1e-3, 2.0, 1e-6)])
def test_safe_interval_is_the_worst_case_limit(f, r, v, dt):
    n = autotune.safe_interval(f, r, v, dt, safety=1.0)
    gap0 = (f - 1.0) * 2 * r          # listed pair furthest from contact
    closed = lambda steps: 2 * v * dt * steps
    assert closed(n) <= gap0 * (1 + 1e-12)
    assert closed(n + 1) > gap0


def test_safety_factor_and_static_case():
    assert autotune.safe_interval(1.2, 0.5e-3, 1.0, 1e-5, safety=2.0) == \
        autotune.safe_interval(1.2, 0.5e-3, 2.0, 1e-5, safety=1.0)
    assert autotune.safe_interval(1.2, 0.5e-3, 0.0, 1e-5) >= 1 << 20
    assert autotune.safe_interval(1.0001, 0.5e-3, 100.0, 1e-5) == 1
