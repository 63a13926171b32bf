"""Ghost cells of U* the step does without (host build of sweeps.hpp).

A run's own steps skip the ghost fill of U* between the sweeps: the x sweep stores
each cell of the interior rows that the four ghost rows copy to those ghost rows
itself (row_push), and the next state's fill reads the cells it copies from the
previous grid through a Dirichlet side from U_old instead of U*.  Both rest on the
closed-form fill (ghost_source, the reference's copy_ghost_cells then L/R then B/T
passes, head/cpu/boundary_cpu.hpp:12-149): these tests check them against it for
every side type and small grids, where the edge rows overlap."""
from __future__ import annotations

import itertools

import pytest

from tests import hostcheck as hc

NG = 2
TYPES = (0, 1, 2, 3)  # periodic, transmissive, reflecting, Dirichlet


@pytest.mark.parametrize("ny", [2, 3, 4, 5, 17])
def test_row_push_matches_the_ghost_fill(ny):
    nx = 5
    for bc in itertools.product(TYPES, repeat=4):
        n, src, dst, flip = hc.row_push(ny, bc[2], bc[3])
        if bc[2] == 3 or bc[3] == 3:
            assert n == 0, bc  # a Dirichlet bottom or top side: the fill runs instead
            continue
        assert n == 4 and sorted(dst) == [0, 1, NG + ny, NG + ny + 1], (bc, dst)
        for e in range(4):
            assert NG <= src[e] < NG + ny
            # the x sweep's pushed row writes, at every interior column, what the fill writes
            for gi in range(NG, NG + nx):
                i, j, prev, fx, fy = hc.ghost_source(gi, dst[e], nx, ny, bc)
                assert (i, j, prev, fx) == (gi, src[e], False, False), (bc, e, gi)
                assert fy == bool((flip >> e) & 1), (bc, e)
            # and only interior rows within two rows of an edge are sources (the sweep's pre-test)
            assert src[e] < 2 * NG or src[e] >= ny


@pytest.mark.parametrize("nx,ny", [(2, 2), (3, 5), (6, 4), (9, 7)])
def test_previous_grid_sources_are_fixed_by_the_intermediate_fill(nx, ny):
    """Every ghost cell the next state's fill copies from the previous grid (U*) is a ghost
    cell that U*'s own fill copies, unflipped, from U_old at the same position: reading
    U_old there gives the same values, whether or not U*'s ghost cells were filled."""
    tnx, tny = nx + 2 * NG, ny + 2 * NG
    ghosts = [(gi, gj) for gj in range(tny) for gi in range(tnx)
              if not (NG <= gi < NG + nx and NG <= gj < NG + ny)]
    checked = 0
    for bc in itertools.product(TYPES, repeat=4):
        for gi, gj in ghosts:  # corners included: the state's ghost cells are all compared
            si, sj, prev, _, _ = hc.ghost_source(gi, gj, nx, ny, bc)
            if not prev:
                continue
            assert not (NG <= si < NG + nx and NG <= sj < NG + ny), (bc, gi, gj)  # always a ghost cell
            i2, j2, prev2, fx2, fy2 = hc.ghost_source(si, sj, nx, ny, bc)
            assert (i2, j2, prev2, fx2, fy2) == (si, sj, True, False, False), (bc, gi, gj, si, sj)
            checked += 1
    assert checked > 0
