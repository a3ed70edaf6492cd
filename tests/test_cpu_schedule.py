"""Launch schedules restated on the CPU: every work item is taken exactly once.

The persistent work-stealing schedule of dac_g-kway_amd/csrc/steal.hpp
(steal_sweeps) and the XCD-aware workgroup order of csrc/dev_common.hpp
(xcd_tile) decide only which wave processes which sweep; the results depend
on every sweep being processed exactly once.  These tests restate both rules
and check that property over many grid shapes, with the queue grabs of the
tail interleaved at random.  The GPU tests check the kernels themselves
(test_persistent_same_result, test_batch_persistent and the parity tests,
whose every kernel runs in XCD order).
"""
import random

import pytest

QUEUES = 8           # kStealQueues
STATIC_WAVES = 2     # kStealStaticWaves


def xcd_tile(b, G):
    per, rem = G // 8, G % 8
    x, i = b % 8, b // 8
    return x * (per + 1) + i if x < rem else rem * (per + 1) + (x - rem) * per + i


@pytest.mark.parametrize("G", list(range(1, 70)) + [255, 256, 257, 1000, 2048, 4097])
def test_xcd_tile_is_a_bijection_in_order_per_xcd(G):
    tiles = [xcd_tile(b, G) for b in range(G)]
    assert sorted(tiles) == list(range(G))
    # Each XCD takes one contiguous run, in increasing dispatch order.
    for x in range(8):
        run = [xcd_tile(b, G) for b in range(x, G, 8)]
        if run:
            assert run == list(range(run[0], run[0] + len(run)))


def steal_schedule(sweeps, grid, wpb, use_queues, rng):
    """Counts how often each sweep is taken under steal_sweeps' rules."""
    W = grid * wpb
    tail = W if (use_queues and sweeps > STATIC_WAVES * W) else 0
    s0 = sweeps - tail
    seen = [0] * sweeps
    for b in range(grid):
        for wave in range(wpb):
            if grid >= QUEUES:
                x, i = b % QUEUES, b // QUEUES
                gx = grid // QUEUES + (1 if x < grid % QUEUES else 0)
                lo, hi = s0 * x // QUEUES, s0 * (x + 1) // QUEUES
                sw = lo + i * wpb + wave
                while sw < hi:
                    seen[sw] += 1
                    sw += gx * wpb
            else:
                sw = b * wpb + wave
                while sw < s0:
                    seen[sw] += 1
                    sw += W
    if tail:
        ctr = [0] * QUEUES

        def workgroup(b):
            home = b % QUEUES
            visits = QUEUES if grid < QUEUES else 2
            for i in range(visits):
                q = (home + i) % QUEUES
                lo, hi = s0 + tail * q // QUEUES, s0 + tail * (q + 1) // QUEUES
                while True:
                    start = ctr[q]          # atomicAdd(c, wpb)
                    ctr[q] += wpb
                    yield
                    if start >= hi - lo:
                        break
                    for wave in range(wpb):
                        sw = lo + start + wave
                        if sw < hi:
                            seen[sw] += 1

        live = [workgroup(b) for b in range(grid)]
        while live:
            g = rng.choice(live)
            try:
                next(g)
            except StopIteration:
                live.remove(g)
    return seen


@pytest.mark.parametrize("seed", range(200))
def test_steal_schedule_takes_every_sweep_once(seed):
    rng = random.Random(seed)
    grid = rng.choice([1, 2, 3, 5, 7, 8, 9, 15, 16, 17, 33, 64, 100])
    wpb = rng.choice([1, 2, 4, 8, 16])
    sweeps = rng.randint(0, 6 * grid * wpb + 50)
    for use_queues in (False, True):
        seen = steal_schedule(sweeps, grid, wpb, use_queues, rng)
        assert all(c == 1 for c in seen), (sweeps, grid, wpb, use_queues)
