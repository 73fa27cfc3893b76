"""Host-side pieces of the ensemble workload (no GPU): the seeded ensemble
definition and the system-aligned block boundaries."""
import numpy as np

from ressim_amd import problems as P
from ressim_amd.ensemble import grid_blocks, system_blocks


def test_config4_sizes():
    prs = P.config4()
    assert len(prs) == 64
    edges = [pr.nx for pr in prs]
    assert all(32 <= e <= 96 for e in edges) and len(set(edges)) > 20
    assert all(pr.nx == pr.ny == pr.nz for pr in prs)
    n = sum(pr.n for pr in prs)
    nnz = sum(P.stencil_nnz(3, pr.nx, pr.ny, pr.nz) for pr in prs)
    assert (n, nnz) == (20988485, 551369773)
    # each system has its own field
    assert not np.array_equal(prs[0].K6[:10], prs[1].K6[:10])


def test_system_blocks():
    ss = np.array([0, 10, 25, 26], np.int32)
    assert system_blocks(ss, 0).tolist() == [0, 10, 25, 26]
    bs = system_blocks(ss, 4)
    assert bs.tolist() == [0, 3, 6, 10, 13, 17, 21, 25, 26]
    assert set(ss.tolist()) <= set(bs.tolist())
    assert np.all(np.diff(bs) > 0) and np.all(np.diff(bs) <= 4)


def test_grid_blocks_are_whole_lines_in_one_plane():
    prs = P.config4(count=3, emin=5, emax=9, seed=3)
    ss = np.concatenate([[0], np.cumsum([pr.n for pr in prs])])
    for rows in (1, 7, 20, 10_000):
        bs = grid_blocks(prs, rows)
        assert bs[0] == 0 and bs[-1] == ss[-1] and np.all(np.diff(bs) > 0)
        assert set(ss.tolist()) <= set(bs.tolist())
        for a, e in zip(bs[:-1], bs[1:]):
            k = np.searchsorted(ss, a, side="right") - 1
            pr, off = prs[k], a - ss[k]
            nx, plane = pr.nx, pr.nx * pr.ny
            assert off % nx == 0 and (e - ss[k]) % nx == 0          # whole lines
            assert off // plane == (e - ss[k] - 1) // plane          # one plane
            assert e - a <= max(rows, nx) and e - a <= plane
