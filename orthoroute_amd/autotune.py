"""Bucket-width selection for the batched SSSP.

The bucket width trades re-expansion (wide buckets: a node's key can drop
several times inside its bucket, and each drop re-expands its row) against
round count (narrow buckets: more round launches, each following its
bucket's chains hop by hop).  Which side dominates depends on the board and
the iteration: on the full 8,000-net backplane the congested iterations'
searches are 5-9x faster at 8,192 buckets than at 128 (128 re-expands every
node many times over), while the late iterations of the 300x300x12 BGA
board are 1.5-2x faster at 128 than at 32,768 (a hundred searches on a
1 M-node lattice: round count dominates).  Paths are identical for every
width; only the time changes.

`base_width` is the width of a batch with no depth estimate (iteration 1):
2.4 typical planar steps.  A batch with a depth estimate (the spread of its
nets' current path costs above their lower bounds) is spread over a target
number of buckets, never narrower than the base width, which `BucketTuner`
chooses online during a negotiation: it measures each batch's search time,
probes the neighbouring targets (x4 up, then down), keeps the fastest,
re-probes every few batches because the costs drift as congestion builds,
and remembers the settled target per lattice shape for the next negotiation
in the process.  A negotiation with nothing settled yet starts at the target
that makes the first batch's buckets one base width wide (the depth over
the base width, to the nearest x4 step): the congestion penalties of a few
edges stretch the depth estimate far beyond the spread of the costs a
search mostly meets, so that a fixed target count starts far too wide on
congested boards (on the 8,000-net backplane 128 buckets took 8x longer
than 8,192 in the first congested iterations).
"""
import math

# targets a tuner moves between (x4 apart); the deep congested searches of
# late iterations on the largest boards settle at 524,288 or more (a 19 M-node
# board with 25,000 nets: iterations 2-8 in 0.18-0.82 s of search at those
# targets against 0.26-2.5 s with 131,072 as the narrowest)
TARGETS = (32, 128, 512, 2048, 8192, 32768, 131072, 524288, 2097152)
_SETTLED = {}   # lattice key -> settled target, for later negotiations in this process


# base width in typical planar steps per search engine: the bucket-round
# engine pays a launch pair per bucket and prefers wide buckets; the one-wave
# engine advances a bucket inside the wave and prefers narrow ones (on the
# 8,000-net backplane pass: 47 ms at 0.15-0.3 cost units, 77 ms at 1.2)
WIDTH_STEPS = {0: 2.4, 1: 0.5}


def base_width(router, cost_scale=0.0, engine=0):
    """Bucket width with no depth estimate: WIDTH_STEPS of the search engine
    times the iteration's median planar cost (cost_scale, rounded up to a
    power of two), or 5/3 of that many base steps when the costs are not
    known yet."""
    k = WIDTH_STEPS.get(int(engine), 2.4)
    if cost_scale > 0.0:
        return max(k * float(cost_scale), 1e-3)
    step = float(router.config.base_cost_weight) * float(router.lattice.pitch)
    return max(step * k * (4.0 / 2.4), 1e-3)


def default_target(n_nodes):
    """Starting target: 128 buckets from 1 M nodes up, 32 below (measured on
    the 600x400x18 backplane and the 96x96x6 board)."""
    return 128 if n_nodes >= 1_000_000 else 32


class BucketTuner:
    """Online choice of the target bucket count for one negotiation.

    The target in use (the incumbent) is measured every batch.  Now and then
    one batch runs at a neighbouring target (x4 up or down, a probe); the
    probe's search time is compared with the incumbent's on the batches just
    before and just after it, so that the drift of the costs as congestion
    builds cancels, and the probe replaces the incumbent when it is at least
    `gain` faster (at once when it takes under `clear` of the incumbent's
    last time; the next probe in the same direction then runs on the very
    next batch).  A failed probe turns the direction; probes follow every
    other batch while they win, then back off (`backoff` batches after two
    failures in a row)."""

    def __init__(self, key, n_nodes, gain=0.25, clear=0.6, backoff=8):
        self.settled = key in _SETTLED
        start = _SETTLED.get(key, default_target(n_nodes))
        self.key = key
        self.inc = TARGETS.index(start) if start in TARGETS else 1
        self.cur = self.inc
        self.started = False
        self.gain, self.clear, self.backoff = float(gain), float(clear), int(backoff)
        self.dir = 1
        self.fails = 0
        self.since = 0
        self.last = None          # incumbent's last time
        self.pending = None       # (probe index, probe time, incumbent time before it)
        self.history = []

    @property
    def best(self):
        return self.inc

    def target(self, ratio=None):
        """The target of the next batch.  ratio: that batch's depth estimate
        over the base width; on the first call of a tuner with no settled
        target for its key, the start moves to the target nearest that ratio
        (log scale), where the buckets are one base width wide."""
        if not self.started:
            self.started = True
            if not self.settled and ratio is not None and ratio > 0.0 and math.isfinite(ratio):
                i = min(range(len(TARGETS)), key=lambda k: abs(math.log(TARGETS[k] / ratio)))
                self.inc = self.cur = i
        return TARGETS[self.cur]

    def _accept(self, i, sec, climb):
        self.inc = i
        self.cur = i
        self.fails = 0
        self.since = 0
        self.last = sec
        _SETTLED[self.key] = TARGETS[i]
        nxt = i + self.dir
        if climb and 0 <= nxt < len(TARGETS):
            self.cur = nxt

    def record(self, seconds):
        """Records the search time of a batch run at target(); picks the
        target of the next batch."""
        sec = float(seconds)
        self.history.append((TARGETS[self.cur], round(sec, 4)))
        if self.cur != self.inc:                       # a probe just ran
            if self.last is not None and sec < self.clear * self.last:
                self._accept(self.cur, sec, climb=True)
                return
            self.pending = (self.cur, sec, self.last)
            self.cur = self.inc
            return
        if self.pending is not None:                   # the batch after a probe
            p, p_sec, before = self.pending
            self.pending = None
            ref = sec if before is None else 0.5 * (before + sec)
            if p_sec < (1.0 - self.gain) * ref:
                self._accept(p, p_sec, climb=False)
                return
            self.fails += 1
            self.dir = -self.dir
        self.last = sec
        self.since += 1
        _SETTLED[self.key] = TARGETS[self.inc]
        wait = 1 if self.fails < 2 else self.backoff
        if self.since >= wait:
            nxt = self.inc + self.dir
            if not 0 <= nxt < len(TARGETS):
                self.dir = -self.dir
                nxt = self.inc + self.dir
            if 0 <= nxt < len(TARGETS):
                self.cur = nxt
                self.since = 0
