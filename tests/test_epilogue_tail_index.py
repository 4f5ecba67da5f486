"""The separate epilogue (k_epilogue, spns_kernels.hip) runs the first
64 floor(nk3 / 64) modes of each (i2, i1) row one wave per row and spreads
the remaining rem = nk3 mod 64 modes of all rows over the grid, row-major:
thread index f -> row = floor((f + 1/2) / rem) in FP64, i3 = nfull + f - row
rem.  This checks that map against integer division for the serving grids,
the configs' grids and rectangular ones: every tail mode once, in range."""
import numpy as np
import pytest

GRIDS = [(32, 32, 32), (64, 64, 64), (96, 64, 64), (128, 128, 128), (192, 128, 128), (256, 256, 256),
         (512, 512, 512), (1024, 1024, 1024), (16, 128, 16), (16, 1024, 16), (32, 32, 16), (64, 32, 48), (48, 48, 48)]


@pytest.mark.parametrize("grid", GRIDS)
def test_tail_map_is_integer_division(grid):
    Nx, Ny, Nz = grid
    nk1, nk2, nk3 = 2 * (Nx // 2) + 1, 2 * (Ny // 2) + 1, Nz // 2 + 1
    nfull = nk3 & ~63
    rem = nk3 - nfull
    if rem == 0:
        pytest.skip("no tail")
    nrow = nk1 * nk2
    f = np.arange(nrow * rem, dtype=np.int64)
    row = ((f.astype(np.float64) + 0.5) * (1.0 / rem)).astype(np.int64)  # the kernel's FP64 form
    assert np.array_equal(row, f // rem)
    i3 = nfull + (f - row * rem)
    assert row.max() == nrow - 1 and i3.min() == nfull and i3.max() == nk3 - 1
    # each (row, i3) of the tail exactly once
    key = row * nk3 + i3
    assert len(np.unique(key)) == len(key)
