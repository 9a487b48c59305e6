"""Graph mode of a context (isar_set_graph_mode): replays of a captured run equal the directly
launched run bit for bit, with eager work between the replays, and new arguments are captured
again (MI355X)."""
from __future__ import annotations

import numpy as np
import pytest

import isar_imaging_cuda_amd as isar
from isar_imaging_cuda_amd import synth

pytestmark = pytest.mark.gpu


# (config, frames): 32 frames of 256 x 4096 are enough echo cells (2^25) for the alignment update's
# two grids per merge level, so both launch shapes are captured and replayed
@pytest.mark.parametrize("cid,K", [(1, 2), (2, 32)], ids=["c1x2", "c2x32"])
def test_graph_replays_match_direct_runs(cid, K):
    import torch
    cfgs = [synth.preset(cid, seed=400 + k) for k in range(min(K, 4))]
    frames = [synth.frame(c) for c in cfgs]
    h = synth.radar_header(cfgs[0])
    M, N = cfgs[0].echo_num, cfgs[0].range_num
    dev = torch.device("cuda:0")
    e = torch.from_numpy(np.stack([frames[k % len(frames)][0] for k in range(K)]).view(np.float32)).to(dev)
    v = torch.from_numpy(np.stack([frames[k % len(frames)][1] for k in range(K)])).to(dev)
    img = torch.zeros((K, M, 512), dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    print(f"[case] graph mode, {K} frames {M}x{N}", flush=True)
    with isar.Imager(isar.RadarParams(M, N, h["band_width"], h["fc"], h["Tp"], h["Fs"]), max_frames=K) as im:
        for opts in (isar.Options(), isar.Options(if_mtrc=False, iterations=40)):
            im.set_graph_mode(False)
            im.run_device(e.data_ptr(), v.data_ptr(), img.data_ptr(), K, opts, stream=s)
            torch.cuda.synchronize()
            direct = img.clone()
            trace_direct = im.trace(1)
            im.set_graph_mode(True)
            for rep in range(3):
                img.zero_()
                im.run_device(e.data_ptr(), v.data_ptr(), img.data_ptr(), K, opts, stream=s)
                torch.cuda.synchronize()
                assert torch.equal(img, direct), f"replay {rep} ({opts})"
                _ = (e[:1, :1024] * 2).sum().item()          # eager work between replays
            t = im.trace(1)
            assert t.cut_offset == trace_direct.cut_offset and np.array_equal(t.phi, trace_direct.phi)
            assert np.array_equal(im.stage("final", 1), direct[1].cpu().numpy())
