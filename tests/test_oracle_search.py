"""The oracle's discrete surrogate search (src/splines.jl:664-850) pinned by hand-derived cases.

- minimal_bounding_box and sorted_corners: the reference's own hand-checked cases (test/splines.jl:404-435),
  restricted to the grid dimension the reference bisects there (the longer one, 1:10), plus the split-point rule.
- bisection_search: the full sequence of evaluated grid indices on a quadratic loss, for which the cubic Hermite
  surrogate is exact, derived by hand from the reference's rules for K = 32 and 64 with nRefAnglesMin = 5.
"""
import numpy as np

GRID10 = np.arange(1.0, 11.0)


def test_minimal_bounding_box_reference_cases(oracle):
    # seen corners {1, 10}: the first split of [1, 10] is at 5
    assert oracle.minimal_bounding_box({1, 10}, GRID10, 2.0) == (1, 5)
    assert oracle.minimal_bounding_box({1, 10}, GRID10, 6.5) == (5, 10)
    # after 5 is seen, both halves are evaluated and split again (at 3 and 7)
    seen = {1, 5, 10}
    assert oracle.minimal_bounding_box(seen, GRID10, 2.0) == (1, 3)
    assert oracle.minimal_bounding_box(seen, GRID10, 4.5) == (3, 5)
    assert oracle.minimal_bounding_box(seen, GRID10, 6.5) == (5, 7)
    assert oracle.minimal_bounding_box(seen, GRID10, 8.5) == (7, 10)


def test_minimal_bounding_box_split_point_and_bottom(oracle):
    # contains() is lo <= x <= hi, and the left half is tested first: x on a split point goes left
    assert oracle.minimal_bounding_box({1, 10}, GRID10, 5.0) == (1, 5)
    assert oracle.minimal_bounding_box({1, 5, 10}, GRID10, 5.0) == (3, 5)
    # a fully evaluated width-1 box is returned (not splittable)
    assert oracle.minimal_bounding_box({1, 2, 3, 5, 10}, GRID10, 2.5) == (2, 3)
    # x outside the left half falls through to the right half
    assert oracle.minimal_bounding_box({1, 10}, GRID10, 10.0) == (5, 10)


def test_sorted_corners_reference_cases(oracle):
    # test/splines.jl:411-414, first coordinate of the box ((2, 4), (3, 7))
    assert oracle.sorted_corners(GRID10, 2, 4, 3.5) == (4, 2)
    assert oracle.sorted_corners(GRID10, 2, 4, 2.5) == (2, 4)
    # a tie keeps the corner order (lo first)
    assert oracle.sorted_corners(GRID10, 2, 4, 3.0) == (2, 4)


def _quadratic(astar):
    return lambda a: ((a - astar) ** 2, 2.0 * (a - astar))


def test_search_sequence_K32(oracle):
    """K = 32 angles in [50, 180] (step 130/31), minimum at 151.3 (between indices 25 and 26).
    initialize! with mineval 5: depth 1 evaluates the corners 1, 32; depth 2 the split 16; depth 3 the splits 8
    and 24 (5 evaluations, stop). bisection_search: the box holding 151.3 is halved through the evaluated boxes
    [16, 32] and [24, 32] to [24, 28] (28 new; 24 is nearer 151.3 but seen); then [24, 26] (26 is nearer: 3.5 vs
    4.9); then [25, 26] (25 new, width 1: converged)."""
    alphas = oracle.linrange(50.0, 180.0, 32)
    x, trace = oracle.surrogate_search(alphas, 5, _quadratic(151.3))
    assert trace == [1, 32, 16, 8, 24, 28, 26, 25]
    assert abs(x - 151.3) < 1e-9


def test_search_sequence_K64(oracle):
    """K = 64 (step 130/63), minimum at 151.3 (between indices 50 and 51). initialize!: 1, 64; 32; 16, 48.
    bisection_search: [48, 56] (56 new), [48, 52] (52 nearer), [50, 52] (151.3 > alpha_50 = 151.11, so the right
    half of [48, 52]; 50 nearer), [50, 51] (51 new, converged)."""
    alphas = oracle.linrange(50.0, 180.0, 64)
    x, trace = oracle.surrogate_search(alphas, 5, _quadratic(151.3))
    assert trace == [1, 64, 32, 16, 48, 56, 52, 50, 51]
    assert abs(x - 151.3) < 1e-9


def test_search_minimum_below_the_grid(oracle):
    """A minimum left of the grid: the surrogate's best point is the first grid angle, so the boxes run down the
    left edge: [1, 8] after initialize! (8 is seen), then [1, 4], [1, 2] (4 and 2 new, 1 nearer but seen)."""
    alphas = oracle.linrange(50.0, 180.0, 32)
    x, trace = oracle.surrogate_search(alphas, 5, _quadratic(20.0))
    assert trace == [1, 32, 16, 8, 24, 4, 2]
    assert x == alphas[0]
