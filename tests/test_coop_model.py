"""CPU model of the barrier schedule of the batch's cooperative launch
(variant 5, k_step_batch_coop in lbm_step_f32.hip).

The kernel runs chunks of `depth` steps; inside a chunk each MALL group runs
the chunk's steps, and a grid-wide barrier follows every step except the
loop's last one and a group's last step of a chunk when there are several
groups and depth >= 2.  A launch that writes the macroscopic fields runs its
last step for every group in a tail after one more barrier.  A cavity's step
s + 1 reads what its step s wrote (across tiles) and overwrites the buffer
that its step s read, so at least one barrier must lie between any two
consecutive steps of the same group.
The model replays the kernel's loop order and checks that for every launch
shape the host can produce (per-launch cap, chunk depth 64 or the whole
launch, one or several groups).
"""
import itertools

import pytest


def schedule(steps, depth, ngroups, macro_last):
    """The kernel's sequence of ('step', group, step) and ('barrier',): the
    chunked loop over the steps, and, when the launch writes the
    macroscopic fields, its last step in a tail after one more barrier."""
    out = []
    nloop = steps - 1 if macro_last else steps
    for c0 in range(0, nloop, depth):
        d = min(depth, nloop - c0)
        for g in range(ngroups):
            for s in range(d):
                step = c0 + s
                out.append(("step", g, step))
                last = step == nloop - 1 and g == ngroups - 1
                group_move = s == d - 1 and ngroups > 1 and d >= 2
                if not last and not group_move:
                    out.append(("barrier",))
    if macro_last:
        if nloop > 0:
            out.append(("barrier",))
        for g in range(ngroups):
            out.append(("step", g, steps - 1))
    return out


def host_launches(nsteps, ngroups, cap=4096, group_depth=64):
    """(steps, depth, macro_last) of each launch that batch_steps makes for
    one call."""
    out = []
    for s0 in range(0, nsteps, cap):
        d = min(cap, nsteps - s0)
        out.append((d, min(group_depth, d) if ngroups > 1 else d, s0 + d == nsteps))
    return out


def check(steps, depth, ngroups, macro_last=False):
    sched = schedule(steps, depth, ngroups, macro_last)
    # every group runs every step once, in order
    for g in range(ngroups):
        assert [e[2] for e in sched if e[0] == "step" and e[1] == g] == list(range(steps))
    # a barrier between any two consecutive steps of one group
    last_pos = {}
    for pos, e in enumerate(sched):
        if e[0] != "step":
            continue
        _, g, s = e
        if s > 0:
            between = sched[last_pos[g] + 1:pos]
            assert ("barrier",) in between, (steps, depth, ngroups, g, s)
        last_pos[g] = pos
    # the launch never ends on a barrier (every workgroup leaves after its
    # last tile), and no two barriers are adjacent
    assert sched[-1][0] == "step"
    for a, b in zip(sched, sched[1:]):
        assert not (a[0] == b[0] == "barrier")
    return sum(e[0] == "barrier" for e in sched)


@pytest.mark.parametrize("nsteps,ngroups", itertools.product(
    [1, 2, 3, 37, 63, 64, 65, 70, 77, 128, 130, 4096, 4100], [1, 2, 3, 6]))
def test_every_consecutive_step_pair_is_separated_by_a_barrier(nsteps, ngroups):
    for steps, depth, macro_last in host_launches(nsteps, ngroups):
        check(steps, depth, ngroups, macro_last)


def test_group_moves_skip_the_barrier_only_when_chunks_hold_one():
    # two groups, chunks of 64: barriers inside chunks only
    assert check(128, 64, 2) == 2 * 2 * 63
    # chunks of one step: every step but the last ends with a barrier
    assert check(5, 1, 2) == 2 * 5 - 1
    # one group: a barrier between every pair of steps
    assert check(100, 100, 1) == 99
    # the macro tail: one barrier before it, none after
    assert check(100, 100, 1, macro_last=True) == 99
    assert check(1, 1, 3, macro_last=True) == 0
    # one loop step per group (depth 1): a barrier after each but the last
    assert check(2, 2, 3, macro_last=True) == 2 + 1
