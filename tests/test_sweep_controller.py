"""The sweep's fallback controller, driven from the host on the CPU: a
scripted batch returns the per-domain residuals a check would return, so the
rules fire (or not) on chosen sequences. Warm starts that turn non-finite or
whose residual grows are reset to the cold start once, with a fresh step
budget; a good start's transient rise and a cold start are never reset;
starts that are not finite or carry another droplet's mass (+10 %) are
screened out before the run."""
import numpy as np

from idmx import sweep
from idmx.sweep import FB_GROWTH, FB_MASS, FB_NONE, FB_NONFINITE, SweepPoint


class ScriptedBatch:
    """Stands in for sweep.Batch: check() returns seq[i](k) for the k-th check
    of live domain i since its (re)start."""

    def __init__(self, seqs, cold):
        self.n = len(seqs)
        self.seqs = list(seqs)
        self.cold = cold
        self.k = np.zeros(self.n, dtype=int)
        self.active = np.ones(self.n, dtype=bool)
        self.ran = False
        self.restarts = []

    def set_active(self, a):
        self.active = np.asarray(a, dtype=bool).copy()

    def run(self, nsteps):
        self.ran = True
        return 0.0

    def check(self):
        out = np.zeros(self.n)
        if not self.ran:          # the reference densities before the first run
            return out
        self.ran = False
        for i in np.nonzero(self.active)[0]:
            out[i] = self.seqs[i](self.k[i])
            self.k[i] += 1
        return out

    def restart(self, i):
        self.restarts.append(i)
        self.seqs[i] = self.cold
        self.k[i] = 0


def decay(a, r):
    return lambda k: a * r ** k


def _relax(seqs, warm, max_checks=60, grow_checks=5, cold=decay(1e-2, 0.3)):
    b = ScriptedBatch(seqs, cold)
    n = len(seqs)
    steps = np.zeros(n, dtype=np.int64)
    reason = np.zeros(n, dtype=int)
    rem, failed, fb, _ = sweep._relax(b, np.ones(n, dtype=bool), np.array(warm), steps, 1e-10,
                                      100, 100 * max_checks, grow_checks, [], b.restart, 't',
                                      reason=reason)
    return b, rem, failed, fb, reason, steps


def test_nan_and_growth_reset_warm_starts_only():
    nan_at_3 = lambda k: np.nan if k >= 3 else 1e-3 * 0.5 ** k
    growing = decay(1e-3, 1.5)
    seqs = [nan_at_3, growing, growing, decay(1e-2, 0.5)]
    b, rem, failed, fb, reason, steps = _relax(seqs, warm=[True, True, False, True])
    assert b.restarts == [0, 1]
    assert list(reason) == [FB_NONFINITE, FB_GROWTH, FB_NONE, FB_NONE]
    assert list(fb) == [True, True, False, False]
    # the restarted domains converge from the cold start; the growing cold
    # domain is never reset and runs out of steps without converging
    assert not rem[0] and not rem[1] and rem[2] and not rem[3]
    assert not failed.any()
    # the growth rule needs grow_checks checks above 10 x best and the start
    g = next(k for k in range(100) if 1e-3 * 1.5 ** k > 1e-2)
    assert steps[1] > 100 * (g + 5 - 1)


def test_transient_rise_of_a_good_start_is_not_reset():
    # falls, rises 15x above its best (still below where it started), then
    # relaxes: the relaxation of a good start is not monotone
    bump = [1e-2, 1e-4, 1.5e-3, 1.4e-3, 1.3e-3, 1.2e-3, 1.1e-3, 1e-3]
    seq = lambda k: bump[k] if k < len(bump) else 1e-3 * 0.4 ** (k - len(bump) + 1)
    b, rem, failed, fb, reason, _ = _relax([seq], warm=[True])
    assert b.restarts == [] and not fb[0] and reason[0] == FB_NONE and not rem[0]


def test_reset_domain_gets_a_fresh_budget():
    # grows for 12 checks (reset there), then the cold start needs 25 checks;
    # the budget is 30 checks per domain since its (re)start
    cold = decay(1e-2, 0.42)
    need = next(k for k in range(200) if 1e-2 * 0.42 ** k < 1e-10)
    assert 20 < need < 30
    b, rem, failed, fb, reason, steps = _relax([decay(1e-3, 1.6)], warm=[True], max_checks=30,
                                               cold=cold)
    assert b.restarts == [0] and reason[0] == FB_GROWTH
    assert not rem[0] and steps[0] > 100 * 30


def test_growth_rule_off():
    b, rem, failed, fb, reason, _ = _relax([decay(1e-3, 1.5)], warm=[True], max_checks=20,
                                           grow_checks=None)
    assert b.restarts == [] and rem[0]


def test_screen_starts_nan_and_mass():
    pts = [SweepPoint(L=32, G=-1.6, R=8.0) for _ in range(4)]
    cold = sweep.cold_start(pts[0])
    starts = [np.full(32 * 32, np.nan), 1.10 * cold, 1.02 * cold, None]
    init, warm, fb, reason, mass0 = sweep.screen_starts(pts, starts, mass_tol=0.05)
    assert list(warm) == [False, False, True, False]
    assert list(fb) == [True, True, False, False]
    assert list(reason) == [FB_NONFINITE, FB_MASS, FB_NONE, FB_NONE]
    assert np.array_equal(init[0], cold) and np.array_equal(init[1], cold)
    assert np.array_equal(init[2], 1.02 * cold)
    assert np.allclose(mass0, cold.sum())


def test_diverging_cold_start_gets_one_gentle_restart():
    """A cold domain that turns non-finite is restarted from the gentle start
    once (fresh budget); one that diverges again is dropped as failed; with
    no gentle restart given it is dropped at once."""
    nan_at = lambda j: (lambda k: np.nan if k >= j else 1e-3)
    for with_gentle in (True, False):
        b = ScriptedBatch([nan_at(2), nan_at(1), decay(1e-2, 0.3)], cold=None)
        gentle_seq = [decay(1e-3, 0.3), nan_at(0), None]
        calls = []

        def gentle(i):
            calls.append(i)
            b.seqs[i] = gentle_seq[i]
            b.k[i] = 0

        n = 3
        steps = np.zeros(n, dtype=np.int64)
        gentled = np.zeros(n, dtype=bool)
        rem, failed, fb, _ = sweep._relax(
            b, np.ones(n, dtype=bool), np.zeros(n, dtype=bool), steps, 1e-10, 100, 6000, 5, [],
            b.restart, 't', restart_gentle=gentle if with_gentle else None, gentled=gentled)
        if with_gentle:
            assert calls == [1, 0] or sorted(calls) == [0, 1]
            assert list(gentled) == [True, True, False]
            assert list(failed) == [False, True, False]
            assert not rem.any()
            assert steps[0] > 300      # its budget restarted: it ran on after the reset
        else:
            assert calls == [] and list(failed) == [True, True, False]
        assert not fb.any() and b.restarts == []
