"""The GPU frame generator (isar_synth_frames_device) against the CPU one (isar_synth_frame):
the same scene, phase-error, walk and noise draws and FP64 sums, so the frames agree sample for
sample except where rounding crosses a float or int16 step (MI355X)."""
from __future__ import annotations

import numpy as np
import pytest

from isar_imaging_cuda_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cid,over", [(1, {}), (4, {}), (5, dict(echo_num=256, range_num=2048)),
                                      (2, dict(snr_db=300.0, peak_adc=0.0, quantize=0))],
                         ids=["c1", "c4", "c5_scene_256x2048", "c2_noise_free"])
def test_gpu_frames_match_cpu_frames(cid, over):
    import torch
    seeds = [11, 12, 13]
    cfg = synth.preset(cid, **over)
    M, N = cfg.echo_num, cfg.range_num
    print(f"[case] GPU synth c{cid} {M}x{N} {over}", flush=True)
    dev = torch.device("cuda:0")
    e = torch.zeros((len(seeds), M, N, 2), dtype=torch.float32, device=dev)
    v = torch.zeros((len(seeds), M), dtype=torch.float64, device=dev)
    synth.frames_device(cfg, seeds, e.data_ptr(), v.data_ptr())
    torch.cuda.synchronize()
    got = e.cpu().numpy().view(np.complex64)[..., 0]
    vel = v.cpu().numpy()
    for k, s in enumerate(seeds):
        cfg.seed = s
        ref, rv = synth.frame(cfg)
        assert np.array_equal(vel[k], rv)
        d = np.abs(got[k] - ref)
        if cfg.quantize:
            # int16 samples: the same value, or one step apart where the FP64 sums straddle .5
            assert d.max() <= 1.5 and np.count_nonzero(d) <= 1e-4 * d.size, (k, d.max(), np.count_nonzero(d))
        else:
            assert np.linalg.norm(got[k] - ref) / np.linalg.norm(ref) < 1e-7
