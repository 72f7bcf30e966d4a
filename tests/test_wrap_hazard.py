"""Why a colour sweep may write its own periodic ghost images only when
every extent of the patch is even.

On a single patch periodic in every dimension, SetGhost after a colour sweep
is the self-periodic wrap (SelfPeriodicBC, LIBRARY/boundary.f90:480-577):
every ghost takes the value of its periodic image.  The sweep kernels fuse
that wrap into the sweep: a thread that updates a boundary cell also writes
the ghosts whose image the cell is.  That is the reference's result only if
no cell of the same colour reads one of those ghosts during the sweep
(LevelRelax, LIBRARY/afasroutines.f90:546-555, sweeps a colour and wraps
afterwards, so every cell of the colour reads the ghosts as they were
before the sweep).

The check below enumerates, for a given extent, the ghosts the colour-c
cells write as images and the ghosts the colour-c cells read through the
5/7-point stencil.  The two sets are disjoint for every even extent and
overlap as soon as one extent is odd (the coarse multigrid levels of a
200 x 36 x 16 root end at 50 x 9 x 4, those of 48 x 24 x 20 at 12 x 6 x 5),
which is why the library fuses the images only for all-even extents and
otherwise sweeps first and wraps after.
"""
import itertools

import pytest


def image_writes(mx, colour):
    """Ghost positions written as periodic images by the colour-`colour`
    cells (the rule of the fused sweep: a cell on the low face writes the
    high ghost and vice versa, edges and corners included)."""
    nd = len(mx)
    out = set()
    for cell in itertools.product(*[range(1, m + 1) for m in mx]):
        if sum(cell) % 2 != colour:
            continue
        choices = []
        for a in range(nd):
            opts = [cell[a]]
            if cell[a] == 1:
                opts.append(mx[a] + 1)
            if cell[a] == mx[a]:
                opts.append(0)
            choices.append(opts)
        for g in itertools.product(*choices):
            if g != cell:
                out.add(g)
    return out


def stencil_ghost_reads(mx, colour):
    """Ghost positions read by the colour-`colour` cells' stencil."""
    nd = len(mx)
    out = set()
    for cell in itertools.product(*[range(1, m + 1) for m in mx]):
        if sum(cell) % 2 != colour:
            continue
        for a in range(nd):
            for d in (-1, 1):
                nb = list(cell)
                nb[a] += d
                if nb[a] == 0 or nb[a] == mx[a] + 1:
                    out.add(tuple(nb))
    return out


@pytest.mark.parametrize("mx", [(8, 8), (16, 6), (8, 8, 8), (12, 6, 10), (4, 2, 2)])
def test_even_extents_images_are_not_read_in_the_same_sweep(mx):
    for c in (0, 1):
        assert not (image_writes(mx, c) & stencil_ghost_reads(mx, c))


@pytest.mark.parametrize("mx", [(7, 8), (50, 9, 4), (12, 6, 5), (9, 9), (5, 4, 4)])
def test_odd_extent_images_race_with_same_colour_reads(mx):
    for c in (0, 1):
        hit = image_writes(mx, c) & stencil_ghost_reads(mx, c)
        assert hit, (mx, c)
        # every such ghost lies on a face normal to an odd extent
        for g in hit:
            assert any((g[a] in (0, mx[a] + 1)) and mx[a] % 2 for a in range(len(mx)))


def test_the_50x9x4_case_named():
    """On 50 x 9 x 4 the colour of cell (i, 1, k) writes ghost (i, 10, k),
    which cell (i, 9, k) of the same colour reads in the same sweep."""
    mx = (50, 9, 4)
    i, k = 2, 1
    c = (i + 1 + k) % 2
    assert (i + 9 + k) % 2 == c
    assert (i, 10, k) in image_writes(mx, c)
    assert (i, 10, k) in stencil_ghost_reads(mx, c)
