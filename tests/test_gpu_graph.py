"""Graph mode of a context (isar_set_graph_mode): replays of a captured run equal the directly
launched run bit for bit, with eager work between the replays, and new arguments are captured
again (MI355X)."""
from __future__ import annotations

import numpy as np
import pytest

import isar_imaging_cuda_amd as isar
from isar_imaging_cuda_amd import synth

pytestmark = pytest.mark.gpu


def test_graph_replays_match_direct_runs():
    import torch
    cfgs = [synth.preset(1, seed=400 + k) for k in range(2)]
    frames = [synth.frame(c) for c in cfgs]
    h = synth.radar_header(cfgs[0])
    M, N = cfgs[0].echo_num, cfgs[0].range_num
    dev = torch.device("cuda:0")
    e = torch.from_numpy(np.stack([f[0] for f in frames]).view(np.float32)).to(dev)
    v = torch.from_numpy(np.stack([f[1] for f in frames])).to(dev)
    img = torch.zeros((2, M, 512), dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    print(f"[case] graph mode, 2 frames {M}x{N}", flush=True)
    with isar.Imager(isar.RadarParams(M, N, h["band_width"], h["fc"], h["Tp"], h["Fs"]), max_frames=2) as im:
        for opts in (isar.Options(), isar.Options(if_mtrc=False, iterations=40)):
            im.set_graph_mode(False)
            im.run_device(e.data_ptr(), v.data_ptr(), img.data_ptr(), 2, opts, stream=s)
            torch.cuda.synchronize()
            direct = img.clone()
            trace_direct = im.trace(1)
            im.set_graph_mode(True)
            for rep in range(3):
                img.zero_()
                im.run_device(e.data_ptr(), v.data_ptr(), img.data_ptr(), 2, opts, stream=s)
                torch.cuda.synchronize()
                assert torch.equal(img, direct), f"replay {rep} ({opts})"
                _ = (e[:1, :1024] * 2).sum().item()          # eager work between replays
            t = im.trace(1)
            assert t.cut_offset == trace_direct.cut_offset and np.array_equal(t.phi, trace_direct.phi)
            assert np.array_equal(im.stage("final", 1), direct[1].cpu().numpy())
