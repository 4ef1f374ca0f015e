"""Host logic of Batch.run_sliced (no GPU): slicing, resume, and the stall test, on a stand-in
for the library calls whose per-case residuals follow prescribed histories.

The stand-in mirrors the batch's documented semantics (include/cfd.h): a run stops a case at
convergence (status 1) or at itc_max (status 4); a resume continues status-4 cases to the
new total; cfd_batch_stop_case stops a case with the caller's status, which a resume keeps.
"""
import numpy as np
import pytest

from cfd_amd import lib as L


class P:
    def __init__(self, imax, jmax):
        self.imax, self.jmax = imax, jmax


class FakeBatch(L.Batch):
    """run() from residual histories res_of[c](step) instead of the GPU."""

    def __init__(self, res_of, tol):
        self.h = None
        self.n = len(res_of)
        self.res_of = res_of
        self.tol = tol
        self.it = np.zeros(self.n, np.int64)
        self.st = np.zeros(self.n, np.int32)
        self.stopped = []
        self.tiles = []                 # (steps done, tiles) per set_tiles call
        self.params = [P(12, 12)] * self.n  # 100 interior cells per case

    def run(self, tol, itc_max, steps_per_graph=50, diverge_factor=0.0, raise_on_fail=False, resume=False):
        if not resume:
            self.it[:] = 0
            self.st[:] = 0
        res = np.zeros((self.n, 3))
        for c in range(self.n):
            if self.st[c] not in (0, 4):
                res[c] = self.res_of[c](self.it[c])
                continue
            while self.it[c] < itc_max:
                self.it[c] += 1
                if self.res_of[c](self.it[c]) < tol and self.it[c] != 1:
                    break
            r = self.res_of[c](self.it[c])
            res[c] = r
            self.st[c] = 1 if r < tol and self.it[c] != 1 else 4
        return 0, self.it.copy(), res, self.st.copy()

    def last_timing(self):
        return 1.0, 1, 1

    def stop_case(self, c, status=5):
        assert status >= 5
        self.stopped.append((c, int(self.it[c])))
        self.st[c] = status

    def set_tiles(self, *t):
        self.tiles.append((int(self.it.max()), t))

    def close(self):
        pass


def decay(rate, floor=0.0):
    """max residual 1 at step 0, falling by a factor `rate` per 10 000 steps down to `floor`."""
    return lambda k: max(rate ** (k / 10000.0), floor)


def first_below(f, tol):
    k = 2
    while not f(k) < tol:
        k += 1
    return k


def test_slices_resume_and_converge():
    fs = [decay(0.1), decay(0.5)]
    b = FakeBatch(fs, tol=1e-3)
    (code, it, res, st), ms = b.run_sliced(1e-3, 200000, slice_steps=10000)
    assert st.tolist() == [1, 1]
    assert it.tolist() == [first_below(f, 1e-3) for f in fs]     # slicing does not change the stop step
    assert not b.stopped


def test_stall_test_stops_only_the_stalled_case():
    # case 0 halves every 10 000 steps; case 1 stalls at 1e-2 after 20 000 steps; case 2 falls
    # by a factor 0.8 per 10 000 steps (0.64 per 20 000 steps: above a 0.5 ratio, a stall)
    b = FakeBatch([decay(0.5), decay(0.1, floor=1e-2), decay(0.8)], tol=1e-6)
    (code, it, res, st), ms = b.run_sliced(1e-6, 300000, slice_steps=10000, stall=(20000, 0.5))
    stopped = dict(b.stopped)
    # the first comparison is at 30 000 steps (against 10 000: a window back from a slice end)
    assert st[0] == 1 and 0 not in stopped and it[0] == first_below(decay(0.5), 1e-6)
    assert st[1] == 5 and stopped[1] == 40000       # 1e-2 at 20 000 and at 40 000
    assert st[2] == 5 and stopped[2] == 30000       # 0.8^3 / 0.8 = 0.64 > 0.5


def test_stall_window_must_be_a_multiple_of_the_slice():
    b = FakeBatch([decay(0.5)], tol=1e-6)
    with pytest.raises(ValueError):
        b.run_sliced(1e-6, 100000, slice_steps=10000, stall=(15000, 0.5))


def test_tail_stages_follow_running_cells():
    # three cases of 100 interior cells: the second converges at step 57 378 (slice end
    # 60 000), the first at 99 658 (slice end 100 000), the third runs to itc_max
    fs = [decay(0.5), decay(0.3), decay(0.95)]
    b = FakeBatch(fs, tol=1e-3)
    stops = [first_below(f, 1e-3) for f in fs[:2]]
    A, B = (16, 16, 1, 1), (8, 8, 1, 1)
    (code, it, res, st), ms = b.run_sliced(1e-3, 300000, slice_steps=10000, tail=[(150, B), (250, A)])
    assert it.tolist() == stops + [300000]          # the switches change no stop step
    assert b.tiles == [(60000, A), (100000, B)]


def test_tail_stage_skipped_when_cells_drop_past_it():
    # both fast cases converge in the same slice: 300 -> 100 running cells, only the last stage
    fs = [decay(0.5), decay(0.5), decay(0.95)]
    b = FakeBatch(fs, tol=1e-3)
    A, B = (16, 16, 1, 1), (8, 8, 1, 1)
    b.run_sliced(1e-3, 300000, slice_steps=10000, tail=[(250, A), (150, B)])
    assert b.tiles == [(100000, B)]


def test_tail_not_set_once_nothing_runs():
    b = FakeBatch([decay(0.5), decay(0.5)], tol=1e-3)
    b.run_sliced(1e-3, 200000, slice_steps=10000, tail=[(150, (8, 8, 1, 1))])
    assert b.tiles == []            # both converge in the same slice: nothing left to run
