"""Element colouring (phc_sem_color_elements): the reference's 4-colour lists
on the SEM2D mesh (SEMmesh.cpp:380-429, restated in the oracle), and in 3D
the 8-colour partition in which no two elements of a colour share a node."""
import ctypes

import numpy as np
import pytest


@pytest.mark.parametrize("ex,ez", [(1, 1), (2, 3), (5, 4), (65, 65), (7, 2)])
def test_2d_lists_are_the_references(phc, oracle, ex, ez):
    mine = phc.sem_color_elements(ex, 0, ez)
    ref = oracle.sem2d_colors(ex, ez)
    assert len(mine) == 4
    for c in range(4):
        assert np.array_equal(mine[c], ref[c]), c
    mpc = ctypes.c_int()
    assert phc.lib().phc_sem_color_count(ex, 0, ez, ctypes.byref(mpc)) == 4
    assert mpc.value == max(len(r) for r in ref)


@pytest.mark.parametrize("ex,ey,ez,p", [(1, 1, 1, 2), (2, 3, 4, 1), (5, 4, 3, 2), (3, 3, 3, 4)])
def test_3d_colours_partition_without_shared_nodes(phc, oracle, ex, ey, ez, p):
    cols = phc.sem_color_elements(ex, ey, ez)
    assert len(cols) == 8
    allel = np.concatenate(cols)
    assert sorted(allel.tolist()) == list(range(ex * ey * ez))
    m = oracle.SemMesh()
    oracle.lib().orc_sem_mesh_init(ctypes.byref(m), ex, ey, ez, 100.0 * ex, 100.0 * ey, 100.0 * ez, p)
    q3 = (p + 1) ** 3
    nodes = np.zeros(ex * ey * ez * q3, np.int32)
    oracle.lib().orc_sem_nodes_list(ctypes.byref(m), nodes)
    nodes = nodes.reshape(-1, q3)
    for c, els in enumerate(cols):
        if len(els) == 0:
            continue
        used = nodes[els].ravel()
        assert len(np.unique(used)) == len(used), c  # no node shared within a colour
        exi, ezi, eyi = els % ex, (els // ex) % ez, els // (ex * ez)
        assert np.all(exi % 2 == (c & 1)) and np.all(ezi % 2 == (c >> 1 & 1)) and np.all(eyi % 2 == (c >> 2 & 1))
        assert np.all(np.diff(els) > 0)  # y, z, x order with x fastest
    # every node is covered by at most one element per colour and by all the
    # elements that hold it across the colours
    count = np.zeros(m.nx * m.ny * m.nz, np.int32)
    for els in cols:
        np.add.at(count, nodes[els].ravel(), 1)
    deg = np.zeros_like(count)
    np.add.at(deg, nodes.ravel(), 1)
    assert np.array_equal(count, deg) and deg.max() <= 8
