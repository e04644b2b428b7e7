"""The kernel-choice cost model (host logic, no device): measured constants in
qubitserf_amd/csrc/tuning.hpp, pass rates scaled by a solve's structure factor. Checked
through qsf_hip_plan_variant on the gross code's Z-side plan and on random plans."""
import math

import numpy as np

from qubitserf_amd import codes, engine
from qubitserf_amd.tune import cdf_below
from tests import oracle as O


def _gross_z():
    Hx, Hz = codes.gross()
    gamma, ranks, check, _ = O.side_plan(1, Hx, Hz)
    return gamma, check


def test_gross_levels_choice():
    """Tight bound 12 on K = 78: the largest levels (2^36+ codewords) take the packed fp4
    test (30 bits, cheaper per codeword than int8) on a fair-coin prior; a structure factor
    as measured on this code (7 at the d = 8 level's 30-bit pass probability) moves d = 8 to
    the 60-bit fp4 test; the small levels stay on the VALU kernel."""
    gamma, check = _gross_z()
    pick = {d: engine.plan_variant(gamma, 144, d, check=check, num_gamma=1, current_best=12) for d in range(3, 11)}
    assert pick[10][3] == 5 and pick[9][3] == 5, pick
    assert all(pick[d][3] == 0 for d in (3, 4, 5, 6)), pick
    structured = engine.plan_variant(gamma, 144, 8, check=check, num_gamma=1, current_best=12, factor=7.0,
                                     at_p=cdf_below(30, 4))
    assert structured[3] == 2, structured


def test_more_structure_never_returns_to_the_narrower_test():
    """At a fixed level, raising the structure factor moves the choice from a 30-bit test
    (int8 or packed fp4) to fp4 (60 bits), never back."""
    gamma, check = _gross_z()
    order = []
    for f in (1, 2, 4, 8, 16, 64, 256):
        v = engine.plan_variant(gamma, 144, 8, check=check, num_gamma=1, current_best=12, factor=f,
                                at_p=cdf_below(30, 4))
        order.append(v[3])
    bits = [{1: 30, 5: 30, 2: 60, 4: 124}[m] for m in order]
    assert bits == sorted(bits) and order[0] in (1, 5) and order[-1] == 2, order


def test_loose_bounds_stay_on_valu():
    """A bound far above the test width's reach (every group would pass) is priced by its
    passing codewords: the VALU kernel with its widest test."""
    rng = np.random.default_rng(5)
    gamma = rng.integers(0, 2**64, size=(1, 90, 3), dtype=np.uint64)
    for r in range(90):
        gamma[0, :, r // 64] &= ~np.uint64(1 << (r % 64))
        gamma[0, r, r // 64] |= np.uint64(1 << (r % 64))
    v = engine.plan_variant(gamma, 192, 4, current_best=40)
    assert v[3] == 0 and v[0] >= 3, v
    tight = engine.plan_variant(gamma, 192, 7, current_best=10)
    assert tight[3] in (1, 5), tight   # a 30-bit matrix test


def test_choice_never_depends_on_device():
    """The choice is a pure function of the plan and the factor (same answer twice)."""
    gamma, check = _gross_z()
    a = [engine.plan_variant(gamma, 144, d, check=check, num_gamma=1, current_best=12) for d in range(3, 11)]
    b = [engine.plan_variant(gamma, 144, d, check=check, num_gamma=1, current_best=12) for d in range(3, 11)]
    assert a == b
    assert math.isfinite(cdf_below(30, 4))
