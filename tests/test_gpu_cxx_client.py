"""The C++ functor adapter (include/artia_kernels.hpp) exercised by a real host program: the
client (tests/cxx_client/client.cpp) calls FPKernel and BPKernel on an fp32 and a half volume on
the GPU, and its outputs equal the same C-ABI calls made through ctypes bit for bit."""
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

import artiatomi_amd as A

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent / "cxx_client"


def client() -> Path:
    exe = HERE / "artia_client"
    if not exe.exists():   # host-only g++ build, seconds
        subprocess.run(["bash", str(HERE / "build.sh")], check=True)
    return exe


def test_cxx_adapter_client_matches_ctypes(gpu, tmp_path):
    nx, ny, nz, W, H, os_, lam = 48, 40, 16, 56, 52, 2, 0.3
    vg = A.volume_geometry(nx, ny, nz)
    tg = A.tilt_geometry(24.0, W, H, psi_deg=1.5, shift=(0.75, -1.25))
    rng = np.random.default_rng(21)
    vol = rng.random((nz, ny, nx), dtype=np.float32)
    volh = rng.random((nz, ny, nx)).astype(np.float16)
    img = rng.standard_normal((H, W)).astype(np.float32)
    (tmp_path / "vg.bin").write_bytes(bytes(vg))
    (tmp_path / "tg.bin").write_bytes(bytes(tg))
    (tmp_path / "params.txt").write_text(f"{W} {H} {os_} {lam!r}\n")
    vol.tofile(tmp_path / "vol_f32.bin")
    volh.view(np.uint16).tofile(tmp_path / "vol_f16.bin")
    img.tofile(tmp_path / "img.bin")
    print(f"running the C++ client on {tmp_path}", flush=True)
    res = subprocess.run([str(client()), str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(res.stdout + res.stderr, flush=True)
    assert res.returncode == 0, res.stderr

    ctx = A.ArtiaContext(0)
    try:
        d_vol, d_volh, d_img = torch.from_numpy(vol).cuda(), torch.from_numpy(volh).cuda(), torch.from_numpy(img).cuda()
        proj, dist = torch.zeros((H, W), device="cuda"), torch.zeros((H, W), device="cuda")
        ctx.fp_march(vg, tg, d_vol, proj, dist)
        proj_h = torch.zeros((H, W), device="cuda")
        ctx.fp_march(vg, tg, d_volh, proj_h, dist)
        ctx.bp(vg, tg, d_vol, d_img, lam, oversampling=os_)
        ctx.bp(vg, tg, d_volh, d_img, lam, oversampling=os_)
        torch.cuda.synchronize()

        def got(name, dtype, shape):
            return np.fromfile(tmp_path / name, dtype=dtype).reshape(shape)

        assert float(np.abs(got("proj_f32.bin", np.float32, (H, W))).max()) > 0
        np.testing.assert_array_equal(got("proj_f32.bin", np.float32, (H, W)), proj.cpu().numpy())
        np.testing.assert_array_equal(got("dist_f32.bin", np.float32, (H, W)), dist.cpu().numpy())
        np.testing.assert_array_equal(got("proj_f16.bin", np.float32, (H, W)), proj_h.cpu().numpy())
        bp32 = got("bp_f32.bin", np.float32, (nz, ny, nx))
        assert not np.array_equal(bp32, vol)
        np.testing.assert_array_equal(bp32, d_vol.cpu().numpy())
        np.testing.assert_array_equal(got("bp_f16.bin", np.uint16, (nz, ny, nx)),
                                      d_volh.cpu().numpy().view(np.uint16))
    finally:
        ctx.close()
