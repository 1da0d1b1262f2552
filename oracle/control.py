"""TEST INFRASTRUCTURE ONLY (the CPU oracle): the run-control rules restated from the reference,
driving any simulation with run(count), dt(), steps() and probe(n) (the oracle's, in tests/).

Simulation::run_until (src/fdtd.rs:1160-1167), Growth (src/fdtd/monitors.rs:665-717),
sources_end (:719-738) and Simulation::run_until_decayed (:740-841). Only tests/ use it, as
the checker.

Pinned by what the rules fix: run_until stops at the first step whose time reaches the target
(its count is the ceiling of the time over dt, checked against integer arithmetic); a pulse in
an open box decays and the run stops at the end of the first block whose probe values all stay
under the fraction of their peak; a continuous wave never decays and the run ends at the limit;
Growth fires only after `needed` rises in a row to a thousand times the level before them
(hand-made level sequences in tests/test_oracle.py).
"""
import math

GROWTH = 1e3


def run_until(sim, time):
    count = 0
    while (sim.steps() + count) * sim.dt() < time:
        count += 1
    sim.run(count)


class Growth:
    def __init__(self, needed):
        self.needed, self.start, self.last, self.rises = needed, 0.0, 0.0, 0

    @staticmethod
    def after(pulse_end, check):
        return Growth(max(math.ceil(pulse_end / check) + 1, 4))

    def check(self, level):
        """The factor the fields grew by, or None while they aren't growing."""
        if self.last > 0.0 and level > self.last:
            if self.rises == 0:
                self.start = self.last
            self.rises += 1
        else:
            self.rises = 0
        self.last = level
        if self.rises >= self.needed and level >= GROWTH * self.start:
            return level / self.start
        return None


def sources_end(waveforms):
    """When the last source ends: a pulse 6 widths after its peak; None with a continuous wave
    (waveforms as (kind, width, delay), kind 2 continuous)."""
    end = 0.0
    for kind, width, delay in waveforms:
        if kind == 2:
            return None
        end = max(end, delay + 6.0 * width)
    return end


class Growing(Exception):
    pass


def run_until_decayed(sim, probes, fraction, interval, limit, waveforms):
    """True if the fields at `probes` decayed; raises Growing if they grow again."""
    end = sources_end(waveforms)
    growth = Growth.after(end, interval) if end is not None else None
    while sim.steps() * sim.dt() < limit:
        start = len(sim.probe(probes[0]))
        began = sim.steps() * sim.dt()
        run_until(sim, min(began + interval, limit))
        sizes = []
        for n in probes:
            v = sim.probe(n)
            peak = max((x * x for x in v), default=0.0)
            recent = max((x * x for x in v[min(start, len(v)):]), default=0.0)
            sizes.append((peak, recent))
        if all(peak > 0.0 and recent <= fraction * peak for peak, recent in sizes):
            return True
        if growth is not None and began >= end:
            factor = growth.check(max(r for _, r in sizes))
            if factor is not None:
                raise Growing(factor)
    return False
