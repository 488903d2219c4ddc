"""The persistent executor's wait bound, replayed on the CPU.

persistent_kernel (csrc/kernels.hip) is disabled: two launches on the device
(n = 256, b = 32, flat tree, 10 s dependency-wait bound) ran past 150 s.  This
model replays its control flow over the real task list (avs_schedule_dump:
dispatch order, per-task dependency counters, signalled tiles) in discrete
time: each workgroup takes the next ticket, polls its dependency counters,
runs the task for one step and increments the counters of the tiles it wrote.
One task's increments are lost, so the tasks that need them can only end by
the wait bound.

Whatever the number of workgroups, the launch ends within about one bound of
the lost signal (each workgroup leaves after its own failed wait), so the
dependency wait cannot account for a run ten times its bound: the stuck point
of the device runs lies elsewhere (the bound's clock and scale were measured
right).  The launch-wide abort the kernel now has (a failed wait sets status;
waits and ticket-taking stop when they see it) only shortens that end, and
runs without a lost signal are unchanged.
"""
from collections import defaultdict

import pytest

from avsvd import schedule_dump


def run_model(tasks, workers, bound, lost=None, abort=True):
    """(end time, failed waits, tasks completed)"""
    n = len(tasks)
    cnt = defaultdict(int)
    ticket, status, failed, completed, time = 0, 0, 0, 0, 0
    state = [("take",)] * workers
    while any(s[0] != "done" for s in state):
        for w in range(workers):
            s = state[w]
            if s[0] == "take":
                if (abort and status) or ticket >= n:
                    state[w] = ("done",)
                    continue
                state[w] = s = ("wait", ticket, time)
                ticket += 1
            if s[0] == "wait":
                _, tk, t0 = s
                if all(cnt[tile] >= need for tile, need in tasks[tk]["deps"]):
                    state[w] = ("run", tk, time + 1)
                elif time - t0 > bound or (abort and status):
                    status = status or tk + 1
                    failed += 1
                    state[w] = ("done",)
            elif s[0] == "run" and time >= s[2]:
                tk = s[1]
                if tk != lost:
                    for tile in tasks[tk]["sigs"]:
                        cnt[tile] += 1
                completed += 1
                state[w] = ("take",)
        time += 1
        assert time < 10 ** 6
    return time, failed, completed


@pytest.fixture(scope="module")
def tasks():
    # the recorded case: n = 256, b = 32 (8 x 8 tiles), Blockbidiag!, flat tree
    return schedule_dump(8, 32, alg=0, kend=0, tree=0, domain=1, workers=16)


@pytest.mark.parametrize("workers", [1, 4, 16, 64, 256])
@pytest.mark.parametrize("where", [4, 2])
def test_a_lost_signal_ends_the_launch_within_about_one_bound(tasks, workers, where):
    bound = 200
    lost = len(tasks) // where
    ok_end = run_model(tasks, workers, bound)[0]
    for abort in (False, True):
        end, failed, completed = run_model(tasks, workers, bound, lost=lost, abort=abort)
        assert 1 <= failed <= workers
        assert completed < len(tasks)
        # the lost task's position, the bound, and at most the normal run's length
        assert end <= bound + ok_end + 2, (abort, end)
    # the abort never lengthens the end
    assert run_model(tasks, workers, bound, lost=lost, abort=True)[0] <= \
        run_model(tasks, workers, bound, lost=lost, abort=False)[0]


@pytest.mark.parametrize("workers", [1, 4, 16])
def test_abort_changes_nothing_when_every_signal_arrives(tasks, workers):
    t_old, f_old, c_old = run_model(tasks, workers, 200, lost=None, abort=False)
    t_new, f_new, c_new = run_model(tasks, workers, 200, lost=None, abort=True)
    assert f_old == f_new == 0
    assert c_old == c_new == len(tasks)
    assert t_old == t_new
