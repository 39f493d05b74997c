"""Host-side geometry helpers: block octrees (nesting closure), batched boxes, rotations, areas."""
import numpy as np

from openaerolab_alpha_01_amd.geometry import (batch_grids, block_octree, ellipsoid_frontal_area, ellipsoid_mask,
                                               random_rotation, sphere_octree, uniform_grid)


def _blocks(g):
    return {tuple(v) for v in (np.asarray(g.coords) - 1).tolist()}


def test_block_octree_children_and_closure():
    grids = block_octree((64, 32, 64), [{(1, 0, 1)}, {(2, 0, 2)}, {(5, 1, 5)}], periodic=(False, False, False))
    for l in range(1, len(grids)):
        fine, coarse = _blocks(grids[l]), _blocks(grids[l - 1])
        nb = [int(v) // 16 for v in grids[l].n]
        parents = {(x // 2, y // 2, z // 2) for x, y, z in fine}
        assert parents <= coarse
        # every parent has all 8 children on the finer level
        for (px, py, pz) in parents:
            for c in range(8):
                assert (2 * px + (c & 1), 2 * py + ((c >> 1) & 1), 2 * pz + (c >> 2)) in fine
        # nesting: a finer block's neighbours (inside the extent) have their parents on this level,
        # so the ghost blocks of the finer level can be filled
        if l >= 2:
            for (x, y, z) in fine:
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            n = (x + dx, y + dy, z + dz)
                            if all(0 <= n[a] < nb[a] for a in range(3)):
                                assert (n[0] // 2, n[1] // 2, n[2] // 2) in coarse


def test_sphere_octree_refines_near_the_surface():
    grids = sphere_octree((64, 32, 32), 3, (128.0, 64.0, 64.0), 12.0, (16, 6))
    assert [g.n for g in grids] == [(64, 32, 32), (128, 64, 64), (256, 128, 128)]
    fine = grids[-1]
    c = fine.cell_coords() + 0.5
    d = np.sqrt(((c - np.array([128.0, 64.0, 64.0])) ** 2).sum(1))
    # the whole sphere and its 6-cell shell lie on the finest level's blocks
    assert d.min() < 1.0
    assert abs(fine.solid.sum() / (4 / 3 * np.pi * 12 ** 3) - 1) < 0.05
    assert fine.n_blocks < (256 // 16) * (128 // 16) ** 2


def test_batch_grids_numbering():
    a, b = uniform_grid((32, 16, 16)), uniform_grid((16, 32, 16))
    big = batch_grids([a, b])
    assert big.n_blocks == a.n_blocks + b.n_blocks
    np.testing.assert_array_equal(big.block_domain, [0] * a.n_blocks + [1] * b.n_blocks)
    nb = np.asarray(big.neighbor)
    # box 0's neighbours stay in box 0, box 1's are offset into box 1
    assert nb[:a.n_blocks].max() <= a.n_blocks
    second = nb[a.n_blocks:]
    assert second[second > 0].min() > a.n_blocks
    np.testing.assert_array_equal(second - np.where(second > 0, a.n_blocks, 0), np.asarray(b.neighbor))


def test_rotations_and_frontal_area():
    rng = np.random.default_rng(3)
    for _ in range(20):
        R = random_rotation(rng)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(R) - 1) < 1e-12
    # sphere: pi r^2 whatever the rotation; axis-aligned ellipsoid: pi b c
    assert abs(ellipsoid_frontal_area((5, 5, 5), random_rotation(rng)) - np.pi * 25) < 1e-10
    assert abs(ellipsoid_frontal_area((3, 4, 5), np.eye(3)) - np.pi * 20) < 1e-10
    # projected area of the voxel mask along x matches the formula to the voxel resolution
    g = uniform_grid((64, 64, 64))
    R = random_rotation(rng)
    m = g.to_dense(ellipsoid_mask(g, (32, 32, 32), (12, 9, 15), R)).astype(bool)
    proj = m.any(axis=0).sum()
    assert abs(proj / ellipsoid_frontal_area((12, 9, 15), R) - 1) < 0.05
