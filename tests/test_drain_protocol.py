"""CPU model of the drain's work-sharing protocol (trace_persistent in
csrc/rt_trace_dev.hpp), checked against the single-lane traversal.

The ordered traversal of one ray visits leaves in a fixed sequence; the
shader's hit_bvh (ray-trace-compute.comp:263-311) keeps the first leaf at the
smallest t (strict `<`).  At any moment a lane's remaining work is a suffix of
that sequence: its current subtree, then its stack entries from top to
bottom.  Taking a lane's bottom stack entry therefore takes the tail segment
of its remaining sequence.  The model keeps exactly that: a lane holds a list
of segments (the first one is in progress), a thief takes the last one, and
the acceptance, piece and hand-over rules are the kernel's:

* a lane holds the piece [plo, plo + 2^pk) of the ray's visiting order;
  a steal halves the donor's piece, the thief gets the second half;
* a lane accepts a leaf strictly closer than its best, or at an equal t when
  its best came from a later piece (best_lo > plo);
* a helper starts from the donor's (best_t, best_lo) with no triangle;
* at hand-over the owner keeps the smaller t, and on equal t the earlier
  piece.

Culling is not modelled: it only removes leaves beyond the current best, so
it never removes a tie.
"""
import math
import random

import pytest


class Lane:
    def __init__(self, segs, plo=0, pk=31, best_t=math.inf, best_lo=0, owner=None):
        self.segs = [list(s) for s in segs if s]
        self.plo, self.pk = plo, pk
        self.best_t, self.best_id, self.best_lo = best_t, -1, best_lo
        self.owner = owner
        self.pending = 0

    def done(self):
        return not self.segs

    def step(self, leaves):
        i = self.segs[0].pop(0)
        if not self.segs[0]:
            self.segs.pop(0)
        t, tid = leaves[i]
        if t is None:
            return
        cap = self.best_t
        if self.best_lo > self.plo:
            cap = math.nextafter(self.best_t, math.inf)
        if t < cap:
            self.best_t, self.best_id, self.best_lo = t, tid, self.plo


def steal(donor, owner, inherit_best_lo=True):
    """The thief takes the donor's bottom entry (its last segment)."""
    assert len(donor.segs) >= 2 and donor.pk > 0
    seg = donor.segs.pop()
    pk = donor.pk - 1
    plo = donor.plo + (1 << pk)
    thief = Lane([seg], plo, pk, donor.best_t, donor.best_lo if inherit_best_lo else plo, owner)
    donor.pk -= 1
    owner.pending += 1
    return thief


def hand_over(helper):
    o = helper.owner
    o.pending -= 1
    if helper.best_id >= 0 and (helper.best_t < o.best_t or
                                (helper.best_t == o.best_t and helper.plo < o.best_lo)):
        o.best_t, o.best_id, o.best_lo = helper.best_t, helper.best_id, helper.plo


def serial(leaves):
    best_t, best = math.inf, -1
    for t, tid in leaves:
        if t is not None and t < best_t:
            best_t, best = t, tid
    return best, best_t


def run_schedule(leaves, segs, rng, inherit_best_lo=True, p_steal=0.3, p_split=0.3):
    """Random interleaving of steps, steals and hand-overs until the owner is done."""
    owner = Lane(segs)
    helpers = []
    while True:
        live = [h for h in helpers if not h.done()]
        finished = [h for h in helpers if h.done()]
        for h in finished:
            hand_over(h)
            helpers.remove(h)
        if owner.done() and owner.pending == 0:
            return owner.best_id, owner.best_t
        # a lane with one long segment may expand it into two (a node whose
        # children are pushed): the remaining order is unchanged
        lanes = [l for l in [owner] + live if not l.done()]
        for l in lanes:
            if rng.random() < p_split and len(l.segs[-1]) >= 2:
                s = l.segs.pop()
                k = rng.randrange(1, len(s))
                l.segs += [s[:k], s[k:]]
        donors = [l for l in lanes if len(l.segs) >= 2 and l.pk > 0]
        if donors and rng.random() < p_steal:
            d = rng.choice(donors)
            helpers.append(steal(d, d.owner or owner, inherit_best_lo))
            continue
        if lanes:
            rng.choice(lanes).step(leaves)


def forced_schedule(inherit_best_lo):
    """Steal, the helper finds t=5 and hands over, steal again, and the second
    helper's piece -- earlier in the visiting order -- holds a leaf at t=5 too."""
    leaves = [(None, 0), (None, 1), (5.0, 2), (None, 3), (5.0, 4)]
    owner = Lane([[0], [1, 2], [3, 4]])
    h1 = steal(owner, owner, inherit_best_lo)        # takes [3, 4]
    while not h1.done():
        h1.step(leaves)
    hand_over(h1)                                     # owner: t=5 from leaf 4
    h2 = steal(owner, owner, inherit_best_lo)        # takes [1, 2]
    while not h2.done():
        h2.step(leaves)
    hand_over(h2)
    while not owner.done():
        owner.step(leaves)
    assert owner.pending == 0
    return owner.best_id, owner.best_t


def test_forced_schedule_equal_t_in_an_earlier_piece():
    want = serial([(None, 0), (None, 1), (5.0, 2), (None, 3), (5.0, 4)])
    assert want == (2, 5.0)
    assert forced_schedule(inherit_best_lo=True) == want


def test_forced_schedule_catches_the_reset_best_lo():
    # a thief that resets best_lo to its own piece rejects the earlier tie
    assert forced_schedule(inherit_best_lo=False) == (4, 5.0)


@pytest.mark.parametrize("seed", range(8))
def test_random_schedules_equal_serial(seed):
    rng = random.Random(seed)
    for _ in range(400):
        m = rng.randrange(1, 40)
        # few distinct t values: ties everywhere
        leaves = [(rng.choice([None, None, 1.0, 2.0, 3.0, 3.0]), i) for i in range(m)]
        cuts = sorted(rng.sample(range(1, m), min(m - 1, rng.randrange(0, 6)))) if m > 1 else []
        segs, lo = [], 0
        for c in cuts + [m]:
            segs.append(list(range(lo, c)))
            lo = c
        want = serial(leaves)
        got = run_schedule(leaves, segs, rng)
        assert got[0] == want[0] and (want[0] < 0 or got[1] == want[1]), (leaves, segs)


def test_random_schedules_find_the_reset_best_lo():
    rng = random.Random(1234)
    bad = 0
    for _ in range(3000):
        m = rng.randrange(2, 24)
        leaves = [(rng.choice([None, None, 1.0, 1.0]), i) for i in range(m)]
        segs = [[i] for i in range(m)]
        if run_schedule(leaves, segs, rng, inherit_best_lo=False, p_steal=0.6)[0] != serial(leaves)[0]:
            bad += 1
    assert bad > 0
