"""The swizzled 8-column LDS layout of the fused z pass (FftS with SWZ,
spns_device.hpp): element (row, b) at (8 row + b) ^ ((row / RPB) mod 8 RPB),
RPB = 4 rows (FP32) or 2 rows (FP64) per 256-byte block.  The transform's
stages do not shift per access: they form the XOR value from the thread's
row and the stage's compile-time offsets.  This checks, for the lengths the
layout is built for, that those shortcut indices equal the direct formula for
every (stage, row, column, r) the stages touch, that the layout is a
bijection onto [0, 8 L), and that the accesses the layout is for spread over
all 64 four-byte banks (no two lanes of a 256-byte access group on one bank
slot).  The GPU parity tests hold the results themselves."""
import numpy as np
import pytest

PLANS = {768: (12, 8, 8), 1536: (12, 16, 8)}


def swz(row, b, rpb):
    sh = 2 if rpb == 4 else 1
    return (8 * row + b) ^ ((row >> sh) & (8 * rpb - 1))


@pytest.mark.parametrize("L", sorted(PLANS))
@pytest.mark.parametrize("rpb", [4, 2])
def test_stage_indices_match_the_layout(L, rpb):
    sh, sm = (2 if rpb == 4 else 1), 8 * rpb - 1
    ns = 1
    for s, R in enumerate(PLANS[L]):
        LR = L // R
        j = np.arange(LR)[:, None]
        b = np.arange(8)[None, :]
        for r in range(R):
            if s > 0:  # reads of rows j + r L/R
                h = ((j >> sh) + ((r * (LR // rpb)) & sm)) & sm
                assert np.array_equal(((8 * j + b) ^ h) + 8 * LR * r, swz(j + r * LR, b, rpb))
            base = (j // ns) * ns * R + j % ns  # writes of rows base + r Ns
            if ns % rpb == 0:
                w = ((8 * base + b) ^ (((base >> sh) + ((r * (ns // rpb)) & sm)) & sm)) + 8 * ns * r
            else:
                assert ns == 1 and R % rpb == 0
                w = 8 * base + ((8 * r + b) ^ (((base >> sh) + (r >> sh)) & sm))
            assert np.array_equal(w, swz(base + r * ns, b, rpb))
        ns *= R
    rows = np.arange(L)[:, None]
    idx = swz(rows, np.arange(8)[None, :], rpb).ravel()
    assert np.array_equal(np.sort(idx), np.arange(8 * L))


def _groups_conflict_free(elem_index, elem_bytes):
    """elem_index: [lanes] element indices of one access; lanes are served in
    groups of 256 bytes (64 banks x 4 B)."""
    per = 256 // elem_bytes
    for g in range(0, len(elem_index), per):
        # first bank of each lane's element; elements are aligned to their size
        slots = (np.asarray(elem_index[g:g + per]) * elem_bytes // 4) % 64
        if len(np.unique(slots)) != len(slots):
            return False
    return True


@pytest.mark.parametrize("rpb,elem_bytes", [(4, 8), (2, 16)])
def test_access_patterns_spread_over_the_banks(rpb, elem_bytes):
    L = 768
    t = np.arange(768)
    # stage reads and the last stage's writes: b = t % 8, rows t // 8 + 96 r
    for r in range(8):
        assert _groups_conflict_free(swz(t // 8 + 96 * r, t % 8, rpb), elem_bytes)
    # the products: one thread per row, one column
    for c in range(8):
        assert _groups_conflict_free(swz(t, c, rpb), elem_bytes)
    # FP32 stage-0 writes with one column per wave: rows 12 j + r across lanes j
    if rpb == 4:
        j = np.arange(64)
        for r in range(12):
            for col in range(8):
                assert _groups_conflict_free(swz(12 * j + r, col, rpb), elem_bytes)
    # the pitch-9 layout it replaced conflicts on the stage reads
    assert not _groups_conflict_free((t // 8 + 96) * 9 + t % 8, elem_bytes)
