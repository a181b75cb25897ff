"""Binomial (revolve-style) checkpoint schedules for the adjoint of an RK solve.

Two levels of storage:
  * snapshots -- full states held by the caller (``snaps`` slots besides the
    initial state);
  * the library's step tape -- ``tape`` slots, each holding one recorded step
    (its five stage states and forward intermediates), so that ``tape``
    consecutive steps are reversed after one recording pass from their start
    state.

R(t, s) is the least number of plain forward steps needed to reverse t steps
whose start state is available, with s free snapshot slots:
    R(t, s) = 0                                   t <= tape
    R(t, 0) = (t - tape) + R(t - tape, 0)         (advance to the last chunk)
    R(t, s) = min_{0<m<t} m + R(t - m, s - 1) + R(m, s)
(advance m steps and keep a snapshot there, reverse the right part with one
slot fewer, then the left part with the slot back).  With tape = 1 this is
Griewank's binomial bound: R(t, s) = r t - C(s + r + 1, s + 2) for the r with
C(s + r, s + 1) < t <= C(s + r + 1, s + 1) (Griewank 1992; revolve, Griewank
and Walther, ACM TOMS 26, 2000), i.e. the optimal count with s + 1 snapshots
counting the initial one.

The forward pass of the solve runs every step anyway, so the snapshots on its
way -- the first split of every nested right part -- cost nothing; F(t, s)
is R with those advances free:
    F(t, s) = 0                                   t <= tape
    F(t, 0) = R(t, 0)
    F(t, s) = min_{0<m<t} F(t - m, s - 1) + R(m, s)
``plan`` minimises F.  With tape = every and snaps = ceil(n / every) - 1 it
gives the uniform schedule (a snapshot every ``every`` steps, no
recomputation beyond the recording passes).
"""
from __future__ import annotations

from functools import lru_cache
from math import comb


class Plan:
    """Optimal splits of an n-step reversal with `snaps` snapshot slots and a
    `tape`-slot step tape."""

    def __init__(self, n: int, snaps: int, tape: int):
        if n < 1 or snaps < 0 or tape < 1:
            raise ValueError("need n >= 1, snaps >= 0, tape >= 1")
        self.n, self.snaps, self.tape = n, snaps, tape

        @lru_cache(maxsize=None)
        def R(t: int, s: int):
            if t <= tape:
                return 0, 0
            if s == 0:
                return (t - tape) + R(t - tape, 0)[0], 0
            best = None
            for m in range(1, t):
                c = m + R(t - m, s - 1)[0] + R(m, s)[0]
                if best is None or c < best[0]:
                    best = (c, m)
            return best

        @lru_cache(maxsize=None)
        def F(t: int, s: int):
            if t <= tape:
                return 0, 0
            if s == 0:
                return R(t, 0)[0], 0
            best = None
            for m in range(1, t):
                c = F(t - m, s - 1)[0] + R(m, s)[0]
                if best is None or c < best[0]:
                    best = (c, m)
            return best

        self.R, self.F = R, F

    def recomputed_steps(self) -> int:
        """Plain forward steps the reversal runs beyond the forward pass."""
        return self.F(self.n, self.snaps)[0]

    def forward_snapshots(self) -> list[int]:
        """Steps after which the forward pass keeps a snapshot (the first
        split of each nested right part)."""
        out, a, s = [], 0, self.snaps
        while self.n - a > self.tape and s > 0:
            m = self.F(self.n - a, s)[1]
            a += m
            out.append(a)
            s -= 1
        return out

    def actions(self):
        """The reversal as a list of actions, given the snapshots of
        forward_snapshots() and the initial state:
          ("restore", a)      state := snapshot at step a
          ("advance", k)      k plain steps
          ("snapshot", a)     keep the current state (step a) in a slot
          ("free", a)         release the slot of step a
          ("reverse", a, k)   record steps a .. a+k-1 on the tape from the
                              current state (step a), then reverse them
        Steps are counted from the start of the solve."""
        acts = []
        fwd = set(self.forward_snapshots())

        def rev(a: int, b: int, s: int, frontier: bool):
            t = b - a
            if t <= self.tape:
                acts.append(("restore", a))
                acts.append(("reverse", a, t))
                return
            if s == 0:
                acts.append(("restore", a))
                acts.append(("advance", t - self.tape))
                acts.append(("reverse", b - self.tape, self.tape))
                rev(a, b - self.tape, 0, False)
                return
            m = (self.F if frontier else self.R)(t, s)[1]
            if not (frontier and a + m in fwd):
                acts.append(("restore", a))
                acts.append(("advance", m))
                acts.append(("snapshot", a + m))
            rev(a + m, b, s - 1, frontier)
            acts.append(("free", a + m))
            rev(a, a + m, s, False)

        rev(0, self.n, self.snaps, True)
        return acts


    def tail(self) -> int:
        """First step of the chunk the reversal takes first, [tail, n): the
        frontier's last chunk, which the forward pass can record on the tape
        as it takes those steps."""
        return next(a[1] for a in self.actions() if a[0] == "reverse")

    def backward_actions(self):
        """actions() for a forward pass that recorded [tail, n) on the tape:
        the actions that only position the state for that recording are
        dropped (the frontier places no snapshot before its last chunk, so
        they are restores and advances), and its reverse becomes
        ("reverse_recorded", tail, k)."""
        acts = self.actions()
        first = next(i for i, a in enumerate(acts) if a[0] == "reverse")
        assert all(a[0] in ("restore", "advance") for a in acts[:first])
        return [("reverse_recorded",) + tuple(acts[first][1:])] + acts[first + 1:]


def binomial_bound(t: int, s: int) -> int:
    """Griewank's least number of forward steps to reverse t steps with s
    snapshots besides the start state and single-step reversal (tape = 1)."""
    if t <= 1:
        return 0
    c = s + 1                      # snapshots counting the start state
    r = 0
    while comb(c + r, c) < t:
        r += 1
    return r * t - comb(c + r, c + 1)
