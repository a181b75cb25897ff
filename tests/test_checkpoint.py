"""Binomial checkpoint schedules of the adjoint (dsem_amd.checkpoint).

* With single-step reversal the optimal recomputation count equals
  Griewank's closed form r t - C(c + r, c + 1) (revolve's bound), for every
  t <= 60 and up to 6 snapshots.
* Played on a model machine, every schedule reverses each step exactly once,
  in reverse order, each from its own start state, never holds more than
  `snaps` snapshots, and advances exactly the planned number of steps.
* tape = every, snaps = n / every - 1 is the uniform schedule: no
  recomputation beyond the recording passes.
* With the forward pass recording the last chunk (Plan.tail), the backward
  actions still reverse every step once from its own state, and record and
  advance no more than the full schedule.
"""
import pytest

from dsem_amd.checkpoint import Plan, binomial_bound


@pytest.mark.parametrize("s", range(0, 7))
def test_single_step_reversal_meets_griewank_bound(s):
    for t in range(1, 61):
        assert Plan(t, s, 1).R(t, s)[0] == binomial_bound(t, s), (t, s)


def play(n, snaps, tape):
    """Run the plan on a machine whose state is the step index."""
    P = Plan(n, snaps, tape)
    held = {0: 0}
    for a in P.forward_snapshots():
        held[a] = a
    assert len(held) - 1 <= snaps
    cur, advanced, reversed_ = None, 0, []
    for act in P.actions():
        if act[0] == "restore":
            cur = held[act[1]]
        elif act[0] == "advance":
            cur += act[1]
            advanced += act[1]
        elif act[0] == "snapshot":
            assert cur == act[1]
            held[act[1]] = cur
            assert len(held) - 1 <= snaps
        elif act[0] == "free":
            del held[act[1]]
        elif act[0] == "reverse":
            a, k = act[1], act[2]
            assert cur == a and 1 <= k <= tape
            reversed_ += list(range(a + k - 1, a - 1, -1))
            cur = None
    assert reversed_ == list(range(n - 1, -1, -1))
    assert advanced == P.recomputed_steps()
    return P


@pytest.mark.parametrize("n,snaps,tape", [(1, 0, 1), (7, 0, 1), (12, 2, 2), (50, 4, 10), (50, 5, 1), (50, 2, 3),
                                          (37, 3, 4), (100, 6, 2)])
def test_schedule_reverses_every_step_once(n, snaps, tape):
    play(n, snaps, tape)


def test_uniform_schedule_is_recovered():
    P = play(50, 4, 10)
    assert P.forward_snapshots() == [10, 20, 30, 40]
    assert P.recomputed_steps() == 0


def test_fewer_snapshots_cost_recomputation():
    costs = [Plan(50, s, 1).recomputed_steps() for s in range(0, 8)]
    assert all(a > b for a, b in zip(costs, costs[1:]))


@pytest.mark.parametrize("n,snaps,tape", [(1, 0, 1), (7, 0, 1), (12, 2, 2), (50, 4, 10), (50, 5, 1), (50, 2, 3),
                                          (37, 3, 4), (100, 6, 2), (50, 0, 10), (10, 1, 5)])
def test_forward_recorded_tail(n, snaps, tape):
    """The machine's forward pass records steps tail .. n-1 on the tape."""
    P = Plan(n, snaps, tape)
    tail = P.tail()
    assert 1 <= n - tail <= tape
    held = {0: 0}
    for a in P.forward_snapshots():
        held[a] = a
    assert all(a <= tail for a in held)
    tape_steps = list(range(tail, n))       # recorded by the forward pass
    cur, advanced, recorded, reversed_ = None, 0, 0, []
    for act in P.backward_actions():
        if act[0] == "restore":
            cur = held[act[1]]
        elif act[0] == "advance":
            cur += act[1]
            advanced += act[1]
        elif act[0] == "snapshot":
            assert cur == act[1]
            held[act[1]] = cur
            assert len(held) - 1 <= snaps
        elif act[0] == "free":
            del held[act[1]]
        elif act[0] in ("reverse", "reverse_recorded"):
            a, k = act[1], act[2]
            if act[0] == "reverse":
                assert cur == a and 1 <= k <= tape
                tape_steps = list(range(a, a + k))
                recorded += k
            assert tape_steps == list(range(a, a + k))
            reversed_ += list(range(a + k - 1, a - 1, -1))
            cur = None
    assert reversed_ == list(range(n - 1, -1, -1))
    assert advanced <= P.recomputed_steps()
    assert recorded == tail
