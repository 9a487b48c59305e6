"""isar_run_device_static: the full run over the library's per-GPU static workspace, with no
device allocation, against isar_run_device on a context (MI355X). Two calls in a row, the
second smaller; a batch beyond the static workspace is refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import isar_imaging_cuda_amd as isar
from isar_imaging_cuda_amd import synth

pytestmark = pytest.mark.gpu


def _batch(cid, K, **over):
    cfgs = [synth.preset(cid, seed=900 + k, **over) for k in range(K)]
    frames = [synth.frame(c) for c in cfgs]
    h = synth.radar_header(cfgs[0])
    p = isar.RadarParams(cfgs[0].echo_num, cfgs[0].range_num, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    return p, np.stack([f[0] for f in frames]), np.stack([f[1] for f in frames])


def test_static_workspace_run_matches_context_run():
    import torch
    dev = torch.device("cuda:0")
    s = torch.cuda.current_stream().cuda_stream
    cases = [("c2, 1 frame", *_batch(2, 1)), ("c1, 2 frames", *_batch(1, 2))]   # the second call is smaller
    bufs = []
    for name, p, echo, vel in cases:
        e = torch.from_numpy(echo.view(np.float32)).to(dev)
        v = torch.from_numpy(vel).to(dev)
        K, M = echo.shape[0], p.echo_num
        img_c = torch.zeros((K, M, 512), dtype=torch.float32, device=dev)
        img_s = torch.full((K, M, 512), float("nan"), dtype=torch.float32, device=dev)
        bufs.append((name, p, K, e, v, img_c, img_s))
    assert isar.static_workspace_bytes() == 64 << 20
    for name, p, K, e, v, img_c, img_s in bufs:
        assert isar.Imager.workspace_bytes(p, K) <= isar.static_workspace_bytes(), name
        print(f"[case] static workspace: {name} {p.echo_num}x{p.range_num}", flush=True)
        with isar.Imager(p, max_frames=K) as im:
            im.run_device(e.data_ptr(), v.data_ptr(), img_c.data_ptr(), K, isar.Options(), stream=s)
            torch.cuda.synchronize()
        isar.run_device_static(p, e.data_ptr(), v.data_ptr(), img_s.data_ptr(), K, isar.Options(), stream=s)
        torch.cuda.synchronize()
        assert torch.equal(img_s, img_c), f"{name}: static run differs from the context run"
    # repeated calls take no device memory (every buffer above already exists)
    name, p, K, e, v, img_c, img_s = bufs[1]
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        img_s.fill_(float("nan"))
        isar.run_device_static(p, e.data_ptr(), v.data_ptr(), img_s.data_ptr(), K, isar.Options(), stream=s)
        torch.cuda.synchronize()
        assert torch.equal(img_s, img_c)
    assert torch.cuda.mem_get_info()[0] == free0


def test_static_workspace_refuses_batches_beyond_it():
    import torch
    p = isar.RadarParams(256, 4096)
    need = isar.Imager.workspace_bytes(p, 8)
    assert need > isar.static_workspace_bytes()
    dev = torch.device("cuda:0")
    img = torch.full((8, 256, 512), 7.0, dtype=torch.float32, device=dev)
    e = torch.zeros((8, 256, 4096, 2), dtype=torch.float32, device=dev)
    v = torch.zeros((8, 256), dtype=torch.float64, device=dev)
    print("[case] static workspace: 8 frames of 256x4096 refused", flush=True)
    with pytest.raises(isar.ImagingError, match="static workspace holds"):
        isar.run_device_static(p, e.data_ptr(), v.data_ptr(), img.data_ptr(), 8, isar.Options())
    torch.cuda.synchronize()
    assert bool((img == 7.0).all())
    with pytest.raises(isar.ImagingError, match="full runs only"):
        isar.run_device_static(p, e.data_ptr(), v.data_ptr(), img.data_ptr(), 1, isar.Options(stage_mode=1))
