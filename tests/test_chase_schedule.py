"""The time schedule of the device chase, checked on the CPU.

avs_bidiagonalize_band_device runs step s of sweep i at time 3i + s, one launch
per time with every active step of that time (api.cpp, chase_store_on_device).
This restates its schedule (the closed-form step count and the active range
[ilo, ihi]) and checks, for many (n, b), what the device relies on: every step
of the host chase (bandbrd.cpp, Chase::steps / Chase::step) runs exactly once,
after the previous step of its own sweep and after the first s+3 steps of the
previous sweep (the host's threaded order), and the row x column footprints of
the steps of one time are pairwise disjoint.
"""
import pytest


def steps_loop(i, b, last):
    # Chase::steps of bandbrd.cpp
    s, r = 1, i + 1
    while min(r + 2 * b - 1, last) > r + b:
        s += 1
        r += b
    return s


def steps_closed(i, b, last):
    x = last - i - 1 - b
    return 1 + ((x + b - 1) // b if x > 0 else 0)


def footprint(i, s, b, last):
    """(rows, cols) ranges step s of sweep i reads or writes (Chase::step)."""
    if s == 0:
        return (i, min(i + 2 * b, last)), (i + 1, min(i + 2 * b, last))
    r = i + 1 + (s - 1) * b
    return (r, min(r + 2 * b - 1, last)), (r + b, min(r + 3 * b - 1, last))


def schedule(n, b):
    """[(t, [(i, s), ...]), ...] as the device driver launches them."""
    last, ns = n - 1, n - 1
    steps = [steps_closed(i, b, last) for i in range(ns)]
    out, ilo, t = [], 0, 0
    while ilo < ns:
        ihi = min(t // 3, ns - 1)
        while ilo <= ihi and t - 3 * ilo >= steps[ilo]:
            ilo += 1
        if ilo <= ihi:
            out.append((t, [(i, t - 3 * i) for i in range(ilo, ihi + 1)]))
        t += 1
    return out, steps


@pytest.mark.parametrize("b", [2, 3, 4, 7, 8, 16, 31, 32, 64, 128])
def test_closed_form_step_count(b):
    for n in (2, 3, b, b + 1, 2 * b, 2 * b + 3, 5 * b + 1, 300):
        for i in range(n - 1):
            assert steps_closed(i, b, n - 1) == steps_loop(i, b, n - 1)


@pytest.mark.parametrize("n,b", [(5, 128), (33, 32), (50, 4), (64, 3), (100, 8), (130, 32), (200, 16),
                                 (257, 2), (300, 128), (520, 64)])
def test_schedule_order_and_disjointness(n, b):
    sched, steps = schedule(n, b)
    when = {}
    for t, act in sched:
        for i, s in act:
            assert 0 <= s < steps[i]
            assert (i, s) not in when
            when[(i, s)] = t
            if s > 0:
                assert when[(i, s - 1)] < t
            if i > 0:
                for q in range(min(s + 3, steps[i - 1])):
                    assert when.get((i - 1, q), float("inf")) < t
        fp = [footprint(i, s, b, n - 1) for i, s in act]
        for x in range(len(fp)):
            for y in range(x + 1, len(fp)):
                (ra, ca), (rb, cb) = fp[x], fp[y]
                rows = not (ra[1] < rb[0] or rb[1] < ra[0])
                cols = not (ca[1] < cb[0] or cb[1] < ca[0])
                assert not (rows and cols), (act[x], act[y])
    assert len(when) == sum(steps)
    # about 3 times per sweep: the launch count the device pays
    assert sched[-1][0] <= 3 * (n - 1) + max(steps)


def test_two_step_lag_would_collide():
    """The lag of 3 is the smallest that works: at a lag of 2, step s of
    sweep i+1 and step s+2 of sweep i share element (r+2b, r+3b)."""
    b, last = 8, 200
    i, s = 10, 2
    (ra, ca) = footprint(i + 1, s, b, last)
    (rb, cb) = footprint(i, s + 2, b, last)
    r = i + 1 + (s - 1) * b
    assert ra[0] <= r + 2 * b <= ra[1] and rb[0] <= r + 2 * b <= rb[1]
    assert ca[0] <= r + 3 * b <= ca[1] and cb[0] <= r + 3 * b <= cb[1]
