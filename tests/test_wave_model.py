"""Host model of the x pass's wave-level run sums (wave_run_sums in csrc/pk_fft.hip).

The kernel combines lanes of a wave that share a bin before touching the LDS histogram:
a segmented inclusive scan built from DPP lane moves, row_shr 1, 2, 4, 8 inside rows of 16
lanes (lanes without a source read 0), with the run flags from wave_shr / wave_shl moves.
This test replays that exact sequence on random slot patterns and checks that the last
lane of every run inside a row holds the run's sum and that summing those lanes equals a
per-bin brute-force histogram.  The full-wave variant (row_bcast15 into rows 1 and 3, then
row_bcast31 into rows 2 and 3) is checked the same way, and a wrong bcast15 row mask is
shown to count a row twice.
"""
import numpy as np
import pytest

W = 64


def dpp(v, ctrl, rows=0xF):
    out = np.zeros_like(v)
    for lane in range(W):
        row = lane // 16
        if not (rows >> row) & 1:
            continue
        src = None
        if ctrl == "wave_shr1":
            src = lane - 1 if lane > 0 else None
        elif ctrl == "wave_shl1":
            src = lane + 1 if lane < W - 1 else None
        elif ctrl[0] == "row_shr":
            k = ctrl[1]
            src = lane - k if lane % 16 >= k else None
        elif ctrl == "bcast15":
            src = (row - 1) * 16 + 15 if row > 0 else None
        elif ctrl == "bcast31":
            src = 31
        if src is not None:
            out[lane] = v[src]
    return out


ROW_STEPS = ((("row_shr", 1), 0xF), (("row_shr", 2), 0xF), (("row_shr", 4), 0xF), (("row_shr", 8), 0xF))
WAVE_STEPS = ROW_STEPS + (("bcast15", 0xA), ("bcast31", 0xC))


def wave_run_sums(slot, vals, rows_only=True):
    """Replays the kernel's sequence; returns (tail mask, scanned values)."""
    lane = np.arange(W)
    prev = dpp(slot + 1, "wave_shr1") - 1
    f = ((lane == 0) | (prev != slot)).astype(np.int64)
    nxt = dpp(f, "wave_shl1") != 0
    tail = (((lane & 15) == 15) if rows_only else (lane == W - 1)) | nxt
    v = vals.astype(np.float64).copy()
    for ctrl, rows in (ROW_STEPS if rows_only else WAVE_STEPS):
        pv = np.stack([dpp(v[:, k], ctrl, rows) for k in range(v.shape[1])], axis=1)
        pf = dpp(f, ctrl, rows)
        upd = f == 0
        v[upd] += pv[upd]
        f = np.where(upd, pf, f)
    return tail, v


@pytest.mark.parametrize("rows_only", [True, False])
@pytest.mark.parametrize("pattern", ["random", "runs", "single", "mostly_invalid", "alternating"])
def test_run_sums_match_brute_force(pattern, rows_only):
    rng = np.random.default_rng(hash(pattern) % 2**32)
    for _ in range(50):
        if pattern == "random":
            slot = rng.integers(-1, 6, W)
        elif pattern == "runs":
            slot = np.sort(rng.integers(-1, 10, W))
            if rng.random() < 0.5:
                slot = slot.reshape(4, 16)[:, ::-1].reshape(-1)   # monotone inside rows only
        elif pattern == "single":
            slot = np.full(W, rng.integers(0, 5))
        elif pattern == "mostly_invalid":
            slot = np.where(rng.random(W) < 0.8, -1, rng.integers(0, 3, W))
        else:
            slot = np.arange(W) % 2
        vals = rng.random((W, 3))
        tail, v = wave_run_sums(slot, vals, rows_only)
        # every run (maximal stretch of equal slot in lane order, cut at row ends when the
        # scan stays inside rows) ends in exactly one tail
        cut = (np.arange(W) % 16 == 0) if rows_only else (np.arange(W) == 0)
        starts = np.r_[True, slot[1:] != slot[:-1]] | cut
        ends = np.r_[starts[1:], True]
        np.testing.assert_array_equal(tail, ends)
        run_id = np.cumsum(starts) - 1
        for r in np.unique(run_id):
            idx = np.nonzero(run_id == r)[0]
            np.testing.assert_allclose(v[idx[-1]], vals[idx].sum(axis=0), rtol=1e-12)
        # histogram from the tails == brute force
        hist = np.zeros((12, 3))
        for lane in np.nonzero(tail & (slot >= 0))[0]:
            hist[slot[lane]] += v[lane]
        want = np.zeros((12, 3))
        for lane in range(W):
            if slot[lane] >= 0:
                want[slot[lane]] += vals[lane]
        np.testing.assert_allclose(hist, want, rtol=1e-12)


def test_wrong_row_mask_would_double_count():
    slot = (np.arange(W) >= 48).astype(np.int64)   # a run ending in row 2 (lane 47)
    vals = np.ones((W, 1))
    lane = np.arange(W)
    f = ((lane == 0) | (lane == 48)).astype(np.int64)
    v = vals.copy()
    for ctrl, rows in ((("row_shr", 1), 0xF), (("row_shr", 2), 0xF), (("row_shr", 4), 0xF), (("row_shr", 8), 0xF),
                       ("bcast15", 0xE), ("bcast31", 0xC)):
        pv = dpp(v[:, 0], ctrl, rows)[:, None]
        pf = dpp(f, ctrl, rows)
        upd = f == 0
        v[upd] += pv[upd]
        f = np.where(upd, pf, f)
    assert v[47, 0] != 48   # the mask 0xe adds row 1 twice into row 2
    tail, good = wave_run_sums(slot, vals, rows_only=False)
    assert good[47, 0] == 48 and good[63, 0] == 16
