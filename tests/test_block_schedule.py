"""Block schedule of the single-grid stage kernel (block_rows in fv2d_amd/csrc/fv2d_stage.h):
every (strip, row) of the grid belongs to exactly one block, for the uniform segments and for
the banded long-then-short schedule; the banded one gives each XCD (block id % 8) one band of
rows, walked top-down, with only short segments in the band's tail and empty blocks only
after a band's segments. The kernel's results do not depend on the schedule (every cell's
arithmetic is the same in any segment); test_emulated_kernel.py checks that bitwise."""
import numpy as np
import pytest

import emu as EMU
from fv2d_amd import problems as P
from fv2d_amd.capi import Params

CASES = [(4096, 4096, 64, 4), (2048, 2048, 32, 4), (1024, 3072, 24, 4), (124, 150, 8, 4), (63, 301, 12, 4),
         (130, 37, 4, 4), (5, 3, 8, 4), (124, 200, 16, 8), (4096, 4096, 64, 8), (62, 1, 4, 4), (700, 9, 8, 8)]


@pytest.fixture(scope="module")
def lib():
    EMU.build()
    return EMU.load()


def coverage(b, Nx, Ny):
    nsx = (Nx + 61) // 62
    cnt = np.zeros((nsx, Ny), dtype=np.int64)
    for sx, jb, je in b:
        assert 0 <= sx < nsx and 0 <= jb <= je <= Ny
        cnt[sx, jb:je] += 1
    return cnt


@pytest.mark.parametrize("banded", [False, True])
@pytest.mark.parametrize("Nx,Ny,seg,nw", CASES)
def test_blocks_tile_the_grid_once(Nx, Ny, seg, nw, banded, lib):
    b = EMU.block_rows(Nx, Ny, seg, nw, banded, lib)
    assert (coverage(b, Nx, Ny) == 1).all()
    if not banded:
        assert (b[:, 2] > b[:, 1]).all()  # no empty block in the uniform schedule


@pytest.mark.parametrize("Nx,Ny,seg,nw", CASES)
def test_banded_order_per_xcd(Nx, Ny, seg, nw, lib):
    b = EMU.block_rows(Nx, Ny, seg, nw, True, lib)
    assert len(b) % 8 == 0
    nchunks = (Ny + nw - 1) // nw
    lseg = (seg + nw - 1) // nw
    sseg = lseg // 4 if lseg >= 8 else min(lseg, 2)
    for x in range(8):
        mine = b[x::8]  # this XCD's blocks in dispatch order
        c0, c1 = x * nchunks // 8, (x + 1) * nchunks // 8
        lo, hi = c0 * nw, min(c1 * nw, Ny)
        live = mine[:, 2] > mine[:, 1]
        # empty (padding) blocks only after the band's segments
        if not live.all():
            assert not live[np.argmin(live):].any()
        seen = {}
        tail0 = max(c1 - lseg, c0) * nw  # first row of the band's tail
        for sx, jb, je in mine[live]:
            assert lo <= jb < je <= hi  # inside the band
            assert jb >= seen.get(sx, -1)  # top-down per strip
            seen[sx] = je
            if jb >= tail0:
                assert je - jb <= sseg * nw  # short segments in the tail
            else:
                assert je - jb <= lseg * nw


def test_kernel_geometry_choice(lib):
    """Large single grids take the banded schedule, small ones the uniform segments."""
    g = EMU.sim_geom(P.kh_params(N=4096), lib)
    assert (g["sched"], g["seg"], g["lseg"], g["sseg"], g["tail"]) == (1, 64, 16, 4, 16)
    assert g["nblocks"] == 8 * 67 * (7 + 4)  # per band: 7 long segments + 4 short ones, 67 strips
    g = EMU.sim_geom(Params.defaults(Nx=2048, Ny=2048), lib)
    assert (g["sched"], g["seg"], g["lseg"], g["sseg"]) == (1, 32, 8, 2)
    g = EMU.sim_geom(Params.defaults(Nx=512, Ny=512), lib)
    assert g["sched"] == 0
    g = EMU.sim_geom(Params.defaults(Nx=4096, Ny=4096, chunk_rows=8), lib)
    assert (g["sched"], g["lseg"], g["sseg"]) == (1, 8, 2)
