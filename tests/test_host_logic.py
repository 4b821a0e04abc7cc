"""Host-side logic that needs no GPU."""
import random

from thruster_cx_code_amd.device import CountBound


def test_count_bound_never_below_true_count():
    """Random growth/removal histories with out-of-order knowledge of past
    counts: the bound always covers the true count and tightens when counts
    arrive."""
    rng = random.Random(0)
    for trial in range(200):
        b = CountBound()
        true = 0
        history = []  # (count after step, snapshot)
        for step in range(60):
            inj = rng.randrange(0, 50)
            b.grow(inj)
            true += inj
            true -= rng.randrange(0, true + 1) if rng.random() < 0.5 else 0
            see = rng.randrange(0, 20)  # possible appends, only some happen
            b.grow(see)
            true += rng.randrange(0, see + 1)
            history.append((true, b.snapshot()))
            # deliver some pending observations, in order, possibly several at once
            while history and rng.random() < 0.4:
                c, snap = history.pop(0)
                b.refresh(c, snap)
            assert b.bound >= true
        while history:
            c, snap = history.pop(0)
            b.refresh(c, snap)
        assert b.bound == true or b.bound >= true


def test_count_bound_tight_after_refresh():
    b = CountBound()
    b.grow(100)
    snap = b.snapshot()
    b.refresh(40, snap)
    assert b.bound == 40
    b.grow(5)
    b.refresh(40, snap)
    assert b.bound == 45


def test_count_bound_with_population_hold():
    """Histories that mix additive growth, removals and top-ups to a target
    (count' = max(count, target)): the bound always covers the true count
    and is at most target + the additive growth not yet observed."""
    rng = random.Random(1)
    for trial in range(200):
        b = CountBound()
        target = rng.randrange(10, 200)
        true = 0
        history = []
        for step in range(60):
            inj = rng.randrange(0, 30)
            b.grow(inj)
            true += inj
            b.top_up(target)
            true = max(true, target)
            true -= rng.randrange(0, true + 1) if rng.random() < 0.7 else 0
            see = rng.randrange(0, 10)
            b.grow(see)
            true += rng.randrange(0, see + 1)
            history.append((true, b.snapshot()))
            while history and rng.random() < 0.5:
                c, snap = history.pop(0)
                b.refresh(c, snap)
            assert b.bound >= true
        while history:
            c, snap = history.pop(0)
            b.refresh(c, snap)
        assert true <= b.bound <= max(true, target)


def test_graph_launch_bounds_cover_the_count_bounds():
    """The batched step's captured graphs are keyed by the count bounds
    rounded up to powers of two: every bound maps to a launch size at least
    as large and less than twice as large (zero stays zero: no launch)."""
    from thruster_cx_code_amd.batch import BatchedSimulator
    p2 = BatchedSimulator._pow2
    assert p2(0) == 0 and p2(-3) == 0 and p2(1) == 1 and p2(2) == 2 and p2(3) == 4
    rng = random.Random(1)
    for _ in range(10000):
        b = rng.randrange(1, 1 << 34)
        q = p2(b)
        assert q >= b and q < 2 * b and q & (q - 1) == 0
