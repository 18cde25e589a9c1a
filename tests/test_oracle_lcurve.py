"""The oracle's L-curve corner search (src/lsqnonneg.jl:872-972) pinned by curves with a known corner.

Curves: P(x) = (log(1 + e^{k(x-c)}), log(1 + e^{-k(x-c)})), an L-curve in log-log form whose Menger curvature peaks
at x = c by symmetry. Each check is one that a plausible mistake fails:
- initial_state and the first move: the first four evaluations are the golden-section points of [-8, 2], and the
  fifth is move_left's (phi x1 + x3)/(phi + 1) when the corner is left of x2 (C2 > C3) and move_right's
  x2 + (x4 - x3) otherwise; a swapped comparison or a mis-typed move formula changes it;
- convergence: the corner is found to within 1e-3 for corners across the bracket (a curvature sign error finds the
  flattest point, a wrong stopping rule stops early);
- the Ctol end-point rule (:881-885): a curve that saturates near mu -> 0 (all points there within 1e-6 of the
  left end point) hides the corner unless those points get -Inf curvature; with Ctol = 0 the search stops at x3;
- backtracking (:892-900): a cached state replaces the current one only if its bracket is no wider, and every move
  shrinks the bracket by 1/phi, so only the current state qualifies; the result is the same with and without it
  (taking the widest or the first cached state instead changes it), and the number of evaluations is bounded by
  4 + ceil(log(10 / xtol) / log(phi)) = 28.
"""
import math

import numpy as np
import pytest

PHI = (1 + 5 ** 0.5) / 2


def soft(c, k=3.0):
    def f(x):
        t = x - c
        return (math.log1p(math.exp(k * t)), math.log1p(math.exp(-k * t)))
    return f


def traced(f):
    calls = []

    def g(x):
        calls.append(x)
        return f(x)
    return g, calls


def test_initial_points_and_first_move(oracle):
    x1, x4 = -8.0, 2.0
    x2 = (PHI * x1 + x4) / (PHI + 1)
    x3 = x1 + (x4 - x2)
    left = (PHI * x1 + x3) / (PHI + 1)   # move_left: corner left of x2
    right = x2 + (x4 - x3)              # move_right
    for c, fifth in ((-6.0, left), (-7.0, left), (0.0, right), (1.0, right)):
        g, calls = traced(soft(c))
        oracle.lcurve_corner(g, x1, x4)
        assert calls[:4] == pytest.approx([x1, x2, x3, x4], abs=0, rel=1e-15)
        assert calls[4] == pytest.approx(fifth, rel=1e-15), c


@pytest.mark.parametrize("c", [-7.0, -6.0, -4.5, -3.0, -1.2, 0.0, 1.0])
def test_corner_recovered(oracle, c):
    x = oracle.lcurve_corner(soft(c), -8.0, 2.0)
    assert abs(x - c) < 1e-3


@pytest.mark.parametrize("c,sat", [(-1.0, -3.5), (0.0, -3.0), (-1.5, -3.8), (0.5, -2.5)])
def test_ctol_end_point_rule(oracle, c, sat):
    B = soft(c, k=0.3)
    Ps = B(sat)

    def f(x):
        if x <= sat:  # saturated: every point within 1e-6 of the left end point
            return (Ps[0] + 1e-6 * math.sin(7 * x), Ps[1] + 1e-6 * math.cos(7 * x))
        return B(x)
    x = oracle.lcurve_corner(f, -8.0, 2.0)
    assert abs(x - c) < 1e-3
    x3 = -8.0 + (2.0 - (PHI * -8.0 + 2.0) / (PHI + 1))
    assert oracle.lcurve_corner(f, -8.0, 2.0, Ctol=0.0) == pytest.approx(x3, rel=1e-12)


def test_backtracking_keeps_the_current_state(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = rng.uniform(-7.5, 1.5, 2)
        ka, kb, wa = rng.uniform(1, 40), rng.uniform(0.5, 5), rng.uniform(0.05, 2)

        def f(x):
            sp = lambda t: math.log1p(math.exp(t)) if t < 30 else t
            return (wa * sp(ka * (x - a)) + sp(kb * (x - b)), wa * sp(-ka * (x - a)) + sp(-kb * (x - b)))
        g, calls = traced(f)
        x_bt = oracle.lcurve_corner(g, -8.0, 2.0, backtracking=True)
        assert len(calls) <= 4 + math.ceil(math.log(10 / 1e-4) / math.log(PHI))
        assert oracle.lcurve_corner(f, -8.0, 2.0, backtracking=False) == x_bt


def test_cache_sizes_on_voxels(dmi, oracle):
    """On real L-curve voxels (the 56-echo, 60-bin, 64-angle shape) the function and point caches never hold more
    than 4 + 24 = 28 entries and the state cache 24: the fixed 32-entry device caches cannot overflow."""
    import ctypes as C

    img, _ = dmi.synth_mse((10, 10, 2), 56, 7.9e-3, ncomp=4, mask_fraction=0.8, seed=5)
    oo = oracle.make_t2map_opts((10, 10, 2), 56, 7.9e-3, 60, (8e-3, 2.0), "lcurve", nRefAngles=64)
    out = (C.c_int * 3)()
    oracle.lib().orc_lcurve_cache_stats(out)  # reset
    oracle.t2map(img, oo, nthreads=8)
    oracle.lib().orc_lcurve_cache_stats(out)
    nfc, npc, nsc = out[0], out[1], out[2]
    assert 4 < nfc <= 28 and 4 < npc <= 28 and nsc <= 24, (nfc, npc, nsc)
