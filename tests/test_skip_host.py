"""The uniform-cube radius that lets the walk cross same-label faces in one step: the
separable restatement the device kernel follows, pinned against a brute-force search."""
import numpy as np
import pytest

from _skip import blob_volume, boundary, skip_radius, skip_radius_brute


@pytest.mark.parametrize("seed,shape,n_labels,cap", [
    (0, (9, 8, 7), 2, 3),
    (1, (12, 10, 9), 3, 15),
    (2, (6, 11, 13), 2, 1),
    (3, (10, 10, 10), 1, 15),       # one label: the radius is the distance to the outer face
])
def test_separable_radius_equals_brute_force(seed, shape, n_labels, cap):
    lab = blob_volume(shape, n_labels, seed)
    assert np.array_equal(skip_radius(lab, cap), skip_radius_brute(lab, cap))


def test_radius_is_a_uniform_cube_and_maximal():
    lab = blob_volume((11, 12, 10), 3, 7)
    r = skip_radius(lab, 15)
    for (x, y, z), k in np.ndenumerate(r):
        k = int(k)
        cube = lab[x - k:x + k + 1, y - k:y + k + 1, z - k:z + k + 1]
        assert cube.shape == (2 * k + 1,) * 3 and np.all(cube == lab[x, y, z])
        j = k + 1
        inside = x - j >= 0 and y - j >= 0 and z - j >= 0 and x + j < 11 and y + j < 12 and z + j < 10
        if inside:
            assert np.any(lab[x - j:x + j + 1, y - j:y + j + 1, z - j:z + j + 1] != lab[x, y, z])


def test_plate_radius_closed_form():
    """A one-label plate in its end-label shell: the radius is the distance to the nearest
    shell-adjacent voxel, min over axes of (distance to the shell - 1), capped."""
    from pymonteopt_amd.geometry import plate_volume
    vol = plate_volume([1.0], 0.1, 1.5, 1.3)
    nx, ny, nz = vol.shape
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    d = np.minimum.reduce([x, nx - 1 - x, y, ny - 1 - y, z, nz - 1 - z])
    want = np.clip(d - 1, 0, 4).astype(np.uint8)
    want[d == 0] = 0
    assert np.array_equal(skip_radius(vol, 4), want)
    assert boundary(vol)[1, 5, 5] and not boundary(vol)[2, 5, 5]


def test_label_255_is_boundary_and_leaves_other_radii():
    """Label 255 (only ever the end label of a 255-label volume) counts as boundary; the
    radii of the other voxels are those of the distance to their nearest other label."""
    lab = blob_volume((14, 12, 13), 2, 5)
    lab[:, :, 9:] = 255
    r = skip_radius(lab, 15)
    assert np.array_equal(r, skip_radius_brute(lab, 15))
    assert np.all(r[lab == 255] == 0)
    other = lab.copy()
    other[other == 255] = 7
    keep = lab != 255
    assert np.array_equal(r[keep], skip_radius(other, 15)[keep])
