"""CPU checks of the CTF SART composition used as the oracle for the GPU CTF driver: the slab loop
covers every ray's whole path through the volume, and the float composition tracks the double
one up to the reference's known float artefacts."""
import numpy as np
import pytest

import artiatomi_amd as A
import ctf_compose as CC
import oracle_lib as O


def _blob(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r2 = ((x - nx / 2 + 0.5) / (0.3 * nx)) ** 2 + ((y - ny / 2 + 0.5) / (0.3 * ny)) ** 2 + \
        ((z - nz / 2 + 0.5) / (0.35 * nz)) ** 2
    return np.exp(-2.0 * r2).astype(np.float32)


@pytest.mark.parametrize("tilt", [0.0, 35.0, -60.0])
def test_slabs_partition_the_ray(tilt):
    nx, ny, nz, W, H = 32, 32, 12, 32, 32
    vol = _blob(nx, ny, nz)
    vg = O.volume_geometry(nx, ny, nz)
    g = O.tilt_geometry(tilt, W, H)
    t_in, t_out = O.defocus_distances(vg, g, W, H)
    total = np.zeros((H, W))
    for ray in CC.slab_rays(t_in, t_out, 3.0):
        total += O.slicer(vg, g, vol, W, H, ray, ray + 3.0, dtype=np.float64)
    full, _ = O.march(vg, g, vol, W, H, dtype=np.float64)
    assert full.max() > 1.0
    # same line integrals; the slabs sample on shifted grids of the same step
    assert np.linalg.norm(total - full) / np.linalg.norm(full) < 0.02


def test_ctf_sart_float_tracks_double():
    nx, ny, nz, W, H = 32, 32, 8, 32, 32
    vol0 = np.zeros((nz, ny, nx), np.float32)
    truth = _blob(nx, ny, nz)
    tilts = [-40.0, -20.0, 0.0, 20.0, 40.0]
    vg = O.volume_geometry(nx, ny, nz)
    pvg = A.volume_geometry(nx, ny, nz)
    acq = A.Acquisition.build(pvg, W, H, tilts)
    projs = np.stack([O.march(vg, g, truth, W, H)[0] for g in acq.geoms])
    ctf = A.ctf_params(pixelsize_nm=0.5, slice_thickness_nm=3.0, beta_fac=(2.0, 0.0, 0.008, 0.0))
    defocus = [(3000.0, 3100.0, 10.0 * i, t) for i, t in enumerate(tilts)]
    sp = A.sart_params(lambda_=0.3)
    args = (vg, vg, acq.geoms, acq.hitpoints, acq.real_lengths, acq.slab_ranges, defocus, acq.index_list, projs,
            vol0, sp, ctf, 2)
    v64, e64 = CC.sart_ctf(*args, dtype=np.float64)
    v32, e32 = CC.sart_ctf(*args, dtype=np.float32)
    assert not e64 and not e32
    assert np.all(np.isfinite(v64)) and np.abs(v64).max() > 1e-3
    # the reference's float order carries the march step bias (float t at |t| ~ 1e4) and the
    # backprojector's chord noise (ray origin 1e5 * n); both stay at the percent level
    d = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    assert 1e-6 < d < 3e-2
