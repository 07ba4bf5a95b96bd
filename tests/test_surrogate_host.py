"""The judged rollout's control flow on the CPU (stand-in calls, no GPU):
kept judged calls, rejection -> revert to the last judged state -> plain
redo -> fallback, and the final validation of the calls after the last
judge (surrogate.SurrogateRollout.run / finish)."""
from __future__ import annotations

from warp_sph_gpt_amd.surrogate import RolloutStats, SurrogateRollout


class Stub(SurrogateRollout):
    """State = list of call kinds applied; the 'surrogate' goes bad at the
    calls listed in `bad` (any state containing a bad call is unhealthy)."""

    def __init__(self, judge_every, fallback_calls, bad=(), slow=()):
        self.judge_every, self.density_tol, self.velocity_tol = judge_every, 0.01, 0.02
        self.fallback_calls = fallback_calls
        self.stats = RolloutStats()
        self._fallback_left = 0
        self._ckpt = None
        self.trace = []
        self.bad = set(bad)
        self.slow = set(slow)       # calls whose velocity error fails the judge
        self.flags = False
        self.insane_end = False

    def _state(self):
        return list(self.trace)

    def _checkpoint(self):
        self._ckpt = (list(self.trace), self.stats.calls)
        self._sur_since = 0

    def _restore(self):
        saved, calls = self._ckpt
        self.trace = list(saved)
        return self.stats.calls - calls

    def surrogate_call(self):
        n = len(self.trace) + 1
        self.trace.append("S!" if n in self.bad else "S")

    def plain_call(self):
        self.trace.append("P")

    def healthy(self):
        return "S!" not in self.trace

    def health(self):
        return {"ok": not self.insane_end or "S" not in self.trace}

    def sane(self, h):
        return h["ok"]

    def _save_start(self):
        return []

    def _restart_plain(self):
        self.trace = []

    def _dem_flags(self):
        return self.flags

    def judged_call(self):
        ok = self.healthy()
        n = len(self.trace) + 1
        self.trace.append("P")
        return {"density_err": 0.001, "velocity_err": 0.5 if n in self.slow else 0.001,
                "state_ok": ok}


def test_clean_run_keeps_every_surrogate_call_and_judges_on_cadence():
    r = Stub(judge_every=5, fallback_calls=3)
    r.run(12)
    r.finish()
    assert r.trace == ["S"] * 4 + ["P"] + ["S"] * 4 + ["P"] + ["S"] * 2
    assert r.stats.judged == 2 and r.stats.rejected == 0 and r.stats.surrogate_calls == 10


def test_bad_interval_is_redone_with_the_plain_step_then_fallback():
    r = Stub(judge_every=5, fallback_calls=3, bad={7})
    r.run(15)
    r.finish()
    # calls 1-4 S, 5 judged P (ckpt); 6-9 S with a bad call -> judge at 10 rejects:
    # back to the state after call 5, calls 6-10 redone plain, 3 fallback calls,
    # 14 S, 15 judged P
    assert r.trace == ["S"] * 4 + ["P"] + ["P"] * 5 + ["P"] * 3 + ["S"] + ["P"]
    assert r.stats.rejected == 1 and r.stats.surrogate_calls == 5
    assert "S!" not in r.trace


def test_calls_after_the_last_judge_are_validated():
    r = Stub(judge_every=5, fallback_calls=2, bad={13})
    r.run(13)
    assert r.trace[-1] == "S!"
    r.finish()
    assert r.trace == ["S"] * 4 + ["P"] + ["S"] * 4 + ["P"] + ["P"] * 3
    assert r.stats.rejected == 1 and r.stats.surrogate_calls == 8


def test_velocity_error_alone_rejects_a_judged_call():
    r = Stub(judge_every=5, fallback_calls=2, slow={5})
    r.run(10)
    r.finish()
    # the judge at call 5 fails on velocity: calls 1-5 redone plain, 2 fallback
    assert r.trace == ["P"] * 5 + ["P"] * 2 + ["S"] * 2 + ["P"]
    assert r.stats.rejected == 1 and r.stats.surrogate_calls == 2


def test_an_insane_end_state_is_redone_from_the_start_with_the_plain_step():
    r = Stub(judge_every=5, fallback_calls=2)
    r.insane_end = True
    r.run(8)
    r.finish()
    assert r.trace == ["P"] * 8
    assert r.stats.restarted_plain and r.stats.surrogate_calls == 0
