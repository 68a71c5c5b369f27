"""Neighbour lists built over bricks of grid points against the FP64 oracle.

atj_grid_count_neighbours / atj_grid_fill_neighbours take a brick of 4 x 8 x 8 grid points per
block and test each point only against the receptor beads the brick can reach.  The lists must
still hold every bead strictly within neighbourdis of the point, in receptor order (the
reference's per-voxel lists, test-calc-grid-energy.py:166-196), padded with 0xFFFF."""
import numpy as np
import pytest

from attract_jax_amd import _lib
from attract_jax_amd.device import DeviceGrid, ForceField, Receptor
from attract_jax_amd.workloads import jittered_receptor, random_pair_table
from oracle import oracle as O

DEV = "cuda:0"


def _check_lists(g, coor, sample):
    nbgrid = g.nbgrid.cpu().numpy()
    nblist = g.nblist.cpu().numpy().view(np.uint16)
    pts = np.asarray(g.origin, np.float64) + g.spacing * np.stack(np.unravel_index(sample, g.dim), 1)
    n_edge = 0
    for v, p in zip(sample, pts):
        count, start = int(nbgrid[v, 0]), int(nbgrid[v, 1])
        got = nblist[start:start + count].astype(np.int64)
        assert np.all(np.diff(got) > 0), "receptor order"
        pad = -(-count // 16) * 16
        assert np.all(nblist[start + count:start + pad] == 0xFFFF)
        ref = O.grid_point_neighbours(p, coor, g.neighbourdis)
        if not np.array_equal(got, ref):  # only beads at the FP32 rounding edge may differ
            dd = np.linalg.norm(coor[np.setxor1d(got, ref)] - p, axis=1)
            assert np.all(np.abs(dd - g.neighbourdis) < 1e-3), dd
            n_edge += 1
    return n_edge


@pytest.mark.gpu
def test_brick_lists_match_the_oracle():
    rng = np.random.default_rng(11)
    coor = jittered_receptor(3000, 25.0 * 1.5 ** (1 / 3), 2.25, rng)
    rc, ac, ivor, _, _ = random_pair_table(rng)
    ff = ForceField.create(rc, ac, ivor, device=DEV)
    rec = Receptor.create(coor, rng.integers(0, 32, 3000), DEV)
    cases = [
        # every bead reachable from every brick: more candidates than one staging pass holds;
        # dims that are no multiple of the brick
        dict(origin=(-20.0, -20.0, -20.0), spacing=2.0, dim=(21, 19, 23), neighbourdis=40.0),
        # most bricks far from every bead, the rest at the serving radius
        dict(origin=(-90.0, -20.0, -20.0), spacing=1.3, dim=(90, 33, 31), neighbourdis=10.8),
    ]
    for kw in cases:
        g = DeviceGrid.build(rec, ff, np.arange(32), device=DEV, potentials=False, plateaudis=10.0,
                             layout=_lib.ATJ_LAYOUT_PLAIN, **kw)
        V = int(np.prod(g.dim))
        # every grid point of the small grid; for the large one the brick faces and corners of
        # the grid plus a random sample
        if V <= 10000:
            sample = np.arange(V)
        else:
            d = np.array(g.dim)
            edge = [np.ravel_multi_index((x, y, z), g.dim) for x in (0, 3, 4, d[0] - 1) for y in (0, 7, 8, d[1] - 1)
                    for z in (0, 7, 8, d[2] - 1)]
            sample = np.unique(np.concatenate([edge, rng.integers(0, V, 1500)]))
        counts = g.nbgrid.cpu().numpy()[:, 0]
        n_edge = _check_lists(g, coor, sample)
        print(f"dim {kw['dim']} neighbourdis {kw['neighbourdis']}: {len(sample)} points checked, "
              f"entries max {counts.max()}, total {counts.sum()}, {n_edge} at the rounding edge", flush=True)
        assert counts.max() > 0
        if kw["neighbourdis"] == 40.0:
            assert counts.max() > 2048  # the staging pass limit of a brick
        else:
            assert (counts == 0).mean() > 0.3  # bricks far from the receptor


@pytest.mark.gpu
def test_shell_order_keeps_every_pair_within_the_plateau():
    """atj_grid_shell_lists reorders each list into distance groups and records, per shell of bead
    positions in the point's cell, how many leading entries a bead there reads.  For random bead
    positions, every receptor bead within plateaudis (FP64) must be among those entries, and each
    reordered list must hold the same entries as before."""
    from attract_jax_amd.workloads import c5_receptor
    r = c5_receptor(("shells", 2000), 2000, np.random.default_rng(2000))
    rng = np.random.default_rng(12)
    rc, ac, ivor, _, _ = random_pair_table(rng)
    ff = ForceField.create(rc, ac, ivor, device=DEV)
    rec = Receptor.create(r.coor, r.types - 1, DEV)
    g = DeviceGrid.build(rec, ff, np.arange(32), device=DEV, potentials=False, layout=_lib.ATJ_LAYOUT_PLAIN,
                         **r.grid_kw)
    whole = g.nblist.cpu().numpy().view(np.uint16).copy()
    nbgrid = g.nbgrid.cpu().numpy()
    nshell = 6
    g.shell_lists(rec, nshell)
    shelled = g.nblist.cpu().numpy().view(np.uint16)
    nbshell = g.nbshell.cpu().numpy().view(np.uint16).reshape(-1, 8)  # start (2 words), c0..c5
    P, h = g.plateaudis, g.spacing * np.sqrt(3.0) / (2 * nshell)
    coor = np.asarray(r.coor, np.float64)
    V = int(np.prod(g.dim))
    counts = nbgrid[:, 0]
    sample = rng.choice(np.nonzero(counts > 0)[0], 300, replace=False)
    checked = 0
    for v in sample:
        n, s = int(nbgrid[v, 0]), int(nbgrid[v, 1])
        assert np.array_equal(np.sort(whole[s:s + n]), np.sort(shelled[s:s + n]))
        p = np.asarray(g.origin, np.float64) + g.spacing * np.array(np.unravel_index(v, g.dim), np.float64)
        for _ in range(8):
            x = p + g.spacing * rng.uniform(-0.5, 0.5, 3)
            k = min(int(np.linalg.norm(x - p) / h), nshell - 1)
            read = set(shelled[s:s + int(nbshell[v, 2 + k])].tolist())
            near = np.nonzero(np.linalg.norm(coor - x, axis=1) < P)[0]
            # neighbourdis covers plateaudis plus the cell's half diagonal: all of them are listed
            assert set(near.tolist()) <= set(whole[s:s + n].tolist()), v
            assert set(near.tolist()) <= read, (v, k)
            checked += len(near)
    assert checked > 1000
