"""The workloads' wall admittances: the hall's branches against the absorption drawn for
each material, the outdoor ground's against Delany-Bazley. Each branch row (m, r, q) is read
back as the continuous admittance 1 / (s M + R + S / s), M = m T, S = q / T (the reference's
branch, impedance.py:168-211, centred in time at fdtd.hpp:391-411), so the checks go through
the rows the solver is given, not through the fit's own parameters."""

import math

import numpy as np
import pytest

from acoustic_spaces_amd import scenes


def _admittance(beta, rows, f, rate):
    s = 2j * math.pi * np.asarray(f, float)
    T = 1.0 / rate
    y = np.full(s.shape, complex(beta))
    for m, r, q in rows:
        y = y + 1.0 / (s * m * T + r + (q / T) / s)
    return y


def _absorption(y):
    return 1.0 - np.abs((1.0 - y) / (1.0 + y)) ** 2


BANDS = (63.0, 125.0, 250.0)


@pytest.mark.parametrize("alpha, miss", [
    ((0.05, 0.12, 0.25), 2e-3), ((0.25, 0.12, 0.05), 2e-3), ((0.1, 0.1, 0.1), 2e-3),
    ((0.05, 0.1, 0.2), 2e-3), ((0.2, 0.1, 0.05), 2e-3),
    # Steeper than first-order branches an octave apart follow, or bent: the closest fit.
    ((0.05, 0.2, 0.3), 0.03), ((0.3, 0.25, 0.1), 0.03), ((0.1, 0.15, 0.12), 0.02), ((0.15, 0.1, 0.15), 0.01)])
def test_branches_reach_an_absorption_that_rises_or_falls(alpha, miss):
    base, rows = scenes.relaxation_branches(alpha, BANDS, 5941.1)
    a = _absorption(_admittance(base, rows, BANDS, 5941.1))
    assert np.max(np.abs(a - alpha)) < miss, a
    assert base >= 0 and np.all(rows[:, 1] > 0) and np.all(rows >= 0)
    # Each row is one kind: rising (m = 0) or falling (q = 0).
    assert np.all((rows[:, 0] == 0) | (rows[:, 2] == 0))
    # The same branches at another rate: the rows change, the admittance does not.
    base2, rows2 = scenes.relaxation_branches(alpha, BANDS, 2 * 5941.1)
    assert np.allclose(_admittance(base2, rows2, BANDS, 2 * 5941.1), _admittance(base, rows, BANDS, 5941.1),
                       rtol=1e-12)


def test_hall_branches_follow_each_materials_drawn_absorption():
    """The hall draws U(0.02, 0.3) at 63, 125 and 250 Hz per material; first-order branches
    an octave apart miss draws that swing up and down by more than they can follow, by
    0.093 at most on this hall."""
    hall = scenes.hall()[0]
    rng = np.random.default_rng(1)
    worst = []
    for k in range(hall.mesh.materials):
        alpha = rng.uniform(0.02, 0.3, 3)
        rows = hall.branches[hall.branch_material == k]
        assert len(rows) == 3
        a = _absorption(_admittance(hall.beta[k], rows, BANDS, hall.rate))
        worst.append(np.max(np.abs(a - alpha)))
        if np.all(np.diff(alpha) > 0) or np.all(np.diff(alpha) < 0):
            assert worst[-1] < 0.05, (k, alpha, a)
    assert max(worst) < 0.1 and np.mean(worst) < 0.05, worst


def test_the_same_alphas_give_the_same_rows():
    a = scenes.relaxation_branches((0.1, 0.2, 0.15), BANDS, 5941.1)
    b = scenes.relaxation_branches(np.array([0.1, 0.2, 0.15]), list(BANDS), 5941.1)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])
    b[1][0, 0] = -1.0  # a copy: the cached fit is not changed through it
    assert np.array_equal(scenes.relaxation_branches((0.1, 0.2, 0.15), BANDS, 5941.1)[1], a[1])


def test_outdoor_ground_is_within_seven_percent_of_delany_bazley():
    o = scenes.outdoor(barrier=False, size=(20.0, 20.0, 8.0))
    rows = o.branches[o.branch_material == 0]
    assert len(rows) == 2 and np.all(rows[:, 0] == 0)
    f = np.geomspace(31.5, 125.0 * math.sqrt(2.0), 9)
    target = scenes.delany_bazley_admittance(f, 200e3)
    y = _admittance(o.beta[0], rows, f, o.rate)
    assert np.max(np.abs(y - target) / np.abs(target)) < 0.07
    # Delany-Bazley at 1 kHz and sigma 200 kPa s/m^2: X = 0.006, Z / (rho c) =
    # 1 + 9.08 X^-0.75 - 11.9j X^-0.73 = 422.18 - 498.31j.
    z = 1.0 / scenes.delany_bazley_admittance(np.array([1000.0]), 200e3)[0]
    assert abs(z - (1 + 9.08 * 0.006 ** -0.75 - 11.9j * 0.006 ** -0.73)) < 1e-9 * abs(z)
    assert abs(z.real - 422.18) < 0.01 and abs(z.imag + 498.31) < 0.01
