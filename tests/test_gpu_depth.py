"""Depth helpers of update_depth on the GPU (wrapper.F90 c_elevation2depth,
c_thickness2vertical_coordinates) against the oracle, bit for bit."""
import numpy as np
import pytest
import torch

from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _grid(nx, ny, nz, halo, seed):
    from getm_amd.core import Grid

    g = Grid(nx, ny, nz, halo, halo, torch.device("cuda", 0))
    rng = np.random.default_rng(seed)
    mask = (rng.random((g.ny_, g.nx_)) > 0.25).astype(np.int32)
    mask[rng.random(mask.shape) < 0.05] = 2  # open-boundary points count as wet
    g.set_field("mask", mask)
    return g, rng, mask


@pytest.mark.parametrize("seed,nx,ny,halo,Dmin", [(1, 37, 29, 2, 0.02), (2, 1, 1, 0, 0.5), (3, 300, 7, 1, -1000.0)])
def test_elevation2depth_bitwise(oracle_mod, seed, nx, ny, halo, Dmin):
    from getm_amd.depth import elevation2depth

    g, rng, mask = _grid(nx, ny, 4, halo, seed)
    shape = mask.shape
    z = rng.uniform(-3.0, 1.0, shape)
    H = rng.uniform(0.0, 4.0, shape)
    mask[0, 0] = 1
    z[0, 0] = -H[0, 0] - 0.1  # dry: clipped to Dmin
    g.set_field("mask", mask)
    D0 = np.full(shape, -7.25)  # kept where the mask is 0
    want = D0.copy()
    oracle_mod.elevation2depth(z, H, Dmin, mask, want)
    d = lambda a: g.wrap(torch.from_numpy(np.ascontiguousarray(a)).cuda())
    D = d(D0)
    elevation2depth(d(z), d(H), Dmin, D)
    torch.cuda.synchronize()
    got = D.numpy()
    assert_bitwise(got, want, f"seed {seed}")
    assert np.all(got[mask == 0] == -7.25)
    if Dmin > 0:
        assert np.any(got[mask != 0] == Dmin), "the case must clip somewhere"


@pytest.mark.parametrize("seed,nx,ny,nz,halo", [(4, 37, 29, 12, 2), (5, 64, 8, 40, 2), (6, 5, 70, 1, 1),
                                                 (7, 130, 3, 2, 0)])
def test_thickness2vertical_coordinates_bitwise(oracle_mod, seed, nx, ny, nz, halo):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.depth import thickness2vertical_coordinates

    g, rng, mask = _grid(nx, ny, nz, halo, seed)
    H = rng.uniform(1.0, 60.0, mask.shape)
    h = rng.uniform(0.05, 3.0, (nz,) + mask.shape)
    zc0 = np.full(h.shape, 1e9)
    zf0 = np.full((nz + 1,) + mask.shape, -1e9)
    zc_w, zf_w = zc0.copy(), zf0.copy()
    oracle_mod.thickness2vertical_coordinates(mask, H, h, zc_w, zf_w)
    d = lambda a, z=None: g.wrap(torch.from_numpy(np.ascontiguousarray(a)).cuda(), z)
    zc, zf = d(zc0, CENTERS), d(zf0, INTERFACES)
    thickness2vertical_coordinates(g.mask, d(H), d(h, CENTERS), zc, zf)
    torch.cuda.synchronize()
    assert_bitwise(zc.numpy(), zc_w, f"zc seed {seed}")
    assert_bitwise(zf.numpy(), zf_w, f"zf seed {seed}")
    dry = mask == 0
    assert np.all(zc.numpy()[:, dry] == 1e9) and np.all(zf.numpy()[:, dry] == -1e9)


def test_depth_helpers_reject_fields_of_the_wrong_kind():
    from getm_amd.core import CENTERS
    from getm_amd.depth import thickness2vertical_coordinates

    g, rng, mask = _grid(8, 6, 3, 1, 8)
    h = g.array(CENTERS, fill=1.0)
    with pytest.raises(ValueError):
        thickness2vertical_coordinates(g.mask, g.array(fill=5.0), h, g.array(CENTERS), g.array(CENTERS))
