"""Device pipeline vs the FP64 oracle, through the C ABI of libisar_imaging.so (MI355X)."""
from __future__ import annotations

import numpy as np
import pytest

import isar_imaging_cuda_amd as isar
from isar_imaging_cuda_amd import synth
from parity import compare, oracle_run

pytestmark = pytest.mark.gpu

FFT_SIZES = [64, 128, 192, 256, 384, 512, 1024, 2048, 4096, 8192, 16384]


@pytest.mark.parametrize("n", FFT_SIZES)
@pytest.mark.parametrize("inverse", [False, True])
def test_device_fft_matches_numpy(n, inverse):
    rng = np.random.default_rng(n + int(inverse))
    x = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    y = isar.fft(x, inverse=inverse)
    ref = np.fft.ifft(x.astype(np.complex128), axis=1) * n if inverse else np.fft.fft(x.astype(np.complex128), axis=1)
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    assert err < 2e-6, f"n={n} inverse={inverse} rel err {err:.2e}"


@pytest.mark.parametrize("n", [64, 128, 192, 256, 384, 512, 1024])
@pytest.mark.parametrize("inverse", [False, True])
def test_wave_ordered_fft_equals_workgroup_ordered(n, inverse):
    """Transforms several per workgroup, each within one wave and ordered by that wave alone (as
    the autofocus, chirp-z and MTRC kernels run them), give the one-transform-per-workgroup
    results bit for bit; a batch that leaves the last workgroup partly empty, and a single row."""
    rng = np.random.default_rng(100 + n + int(inverse))
    for rows in (1, 37):
        x = (rng.standard_normal((rows, n)) + 1j * rng.standard_normal((rows, n))).astype(np.complex64)
        print(f"[case] n={n} inverse={inverse} rows={rows}", flush=True)
        assert np.array_equal(isar.fft(x, inverse=inverse, grouped=True), isar.fft(x, inverse=inverse))


def _frame(config_id, **over):
    cfg = synth.preset(config_id, **over)
    echo, vel = synth.frame(cfg)
    return cfg, echo, vel, synth.radar_header(cfg)


def _run_device(echo, vel, radar, M, N, opts: isar.Options, frames=1):
    params = isar.RadarParams(M, N, radar["band_width"], radar["fc"], radar["Tp"], radar["Fs"])
    with isar.Imager(params, max_frames=frames) as im:
        img = im.run_host(echo, vel, opts)
        traces = [im.trace(k) for k in range(frames)]
    return img, traces


CASES = [
    ("c1_M64", 1, dict(echo_num=64)),
    ("c1", 1, {}),
    ("c1_M192", 1, dict(echo_num=192)),     # pulse counts that are not powers of two
    ("c1_M384", 1, dict(echo_num=384)),
    ("c1_M512", 1, dict(echo_num=512)),
    ("c2", 2, {}),
    ("c4", 4, {}),                          # 256 x 2048, range-aligned scene
    ("c1_N8192", 1, dict(range_num=8192)),   # the 8192-cell range transforms
    ("c1_M256_N8192", 1, dict(echo_num=256, range_num=8192)),
    ("c5", 5, {}),                          # 1024 x 16384, 1000 scatterers, 7 km/s
]

# Batches of copies of a case's frame large enough (frames x pulses x range cells >= 2^25) for
# the alignment update's two-grid launch; every copy must equal the single frame bit for bit.
COPIES = {"c1": 256, "c4": 64, "c2": 32, "c5": 2}



@pytest.mark.parametrize("name,cid,over", CASES, ids=[c[0] for c in CASES])
def test_pipeline_parity(name, cid, over):
    cfg, echo, vel, radar = _frame(cid, **over)
    M, N = cfg.echo_num, cfg.range_num
    print(f"[case] {name} M={M} N={N}", flush=True)
    img, (t,) = _run_device(echo, vel, radar, M, N, isar.Options())
    img_o, tr = oracle_run(echo, vel, radar, M, N)
    rep = compare(img, t, img_o, tr, M)
    print(rep, f"max lag error {np.abs(np.asarray(t.ra_lag) - tr.ra_lag[:M - 1]).max():.2e} cell")
    assert rep.ok(), str(rep)
    n = COPIES.get(name)
    if n:
        print(f"[case] {name}: {n} copies in one batch", flush=True)
        imgs, ts = _run_device(np.stack([echo] * n), np.stack([vel] * n), radar, M, N, isar.Options(), frames=n)
        for k in range(n):
            assert np.array_equal(imgs[k], img), f"copy {k} differs from the single frame"
            assert np.array_equal(ts[k].phi, t.phi) and ts[k].cut_offset == t.cut_offset


@pytest.mark.parametrize("cid,seeds", [(1, range(500, 548)), (2, range(600, 616)), (4, range(700, 716))],
                         ids=["c1x48", "c2x16", "c4x16"])
def test_pipeline_parity_seed_sweep(cid, seeds):
    """One batch of frames with different seeds (scene, motion, phase error and noise draws),
    every frame against the oracle: the batched path over many draws, near-tie decisions
    included."""
    cfgs = [synth.preset(cid, seed=s) for s in seeds]
    frames = [synth.frame(c) for c in cfgs]
    radar = synth.radar_header(cfgs[0])
    M, N = cfgs[0].echo_num, cfgs[0].range_num
    K = len(cfgs)
    print(f"[case] seed sweep c{cid}: {K} frames {M}x{N} in one batch", flush=True)
    echo = np.stack([f[0] for f in frames])
    vel = np.stack([f[1] for f in frames])
    img, traces = _run_device(echo, vel, radar, M, N, isar.Options(), frames=K)
    bad = []
    for k in range(K):
        img_o, tr = oracle_run(echo[k], vel[k], radar, M, N)
        rep = compare(img[k], traces[k], img_o, tr, M)
        if not rep.ok():
            bad.append((k, str(rep)))
    assert not bad, bad


def test_degenerate_frames_match_oracle():
    """An all-zero frame, complex noise with no target, one beat tone in every pulse, and a real
    frame scaled by 1e4, by 1e-4 and by 1e12 (past the 32-bit wrap of the selection threshold's
    int sum), imaged as one batch: no NaN, and the oracle's decisions and
    images (the zero frame gives a zero image and no centring). The tone's spectrum is symmetric
    about its peak, so its centring and bin selection tie exactly in the oracle; those are
    reported as near ties and its image is held to the usual bar."""
    cfg = synth.preset(1, seed=321)
    real, vel = synth.frame(cfg)
    radar = synth.radar_header(cfg)
    M, N = cfg.echo_num, cfg.range_num
    rng = np.random.default_rng(5)
    noise = (100 * (rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N)))).astype(np.complex64)
    tone = (1000 * np.tile(np.exp(2j * np.pi * np.arange(N) * 137 / N), (M, 1))).astype(np.complex64)
    frames = [np.zeros((M, N), np.complex64), noise, tone, (real * 1e4).astype(np.complex64),
              (real * 1e-4).astype(np.complex64), (real * 1e12).astype(np.complex64)]
    vels = [np.zeros(M), np.zeros(M), np.zeros(M), vel, vel, vel]
    print(f"[case] degenerate frames, batch of {len(frames)}", flush=True)
    img, traces = _run_device(np.stack(frames), np.stack(vels), radar, M, N, isar.Options(), frames=len(frames))
    assert np.isfinite(img).all()
    assert not img[0].any() and traces[0].center_applied == 0
    for k in range(len(frames)):
        img_o, tr = oracle_run(frames[k], vels[k], radar, M, N)
        rep = compare(img[k], traces[k], img_o, tr, M)
        assert rep.ok(), f"frame {k}: {rep}"


VARIANTS = [
    ("no_mtrc", dict(if_mtrc=False), dict(if_mtrc=0)),
    ("window_intended", dict(quirk_q1=False), dict(quirk_q1=0)),
    ("hrrp", dict(if_hrrp=True), dict(if_hrrp=1)),
    ("no_hpc", dict(if_hpc=False), dict(if_hpc=0)),
    ("ifds", dict(data_type=1), dict(data_type=1)),
]


@pytest.mark.parametrize("name,dev,orc", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_pipeline_variants(name, dev, orc):
    cfg, echo, vel, radar = _frame(1)
    M, N = cfg.echo_num, cfg.range_num
    print(f"[case] variant {name}", flush=True)
    img, (t,) = _run_device(echo, vel, radar, M, N, isar.Options(**dev))
    img_o, tr = oracle_run(echo, vel, radar, M, N, **orc)
    rep = compare(img, t, img_o, tr, M)
    print(rep)
    assert rep.ok(), str(rep)


@pytest.mark.parametrize("M,N", [(64, 1024), (128, 1024), (192, 1024), (256, 2048), (384, 1024), (512, 1024),
                                 (1024, 2048)])
def test_autofocus_forms_agree(M, N):
    """One workgroup per frame, the split forms (several workgroups per frame, one or two launches
    per iteration) and the resident form (several workgroups per frame exchanging their sums
    inside one launch) give the same phases and images bit for bit; no frame reports a resident
    wait that hit its bound (status bit 2)."""
    cfgs = [synth.preset(1, echo_num=M, range_num=N, seed=300 + k) for k in range(2)]
    frames = [synth.frame(c) for c in cfgs]
    radar = synth.radar_header(cfgs[0])
    echo = np.stack([f[0] for f in frames])
    vel = np.stack([f[1] for f in frames])
    print(f"[case] autofocus forms M={M} N={N}", flush=True)
    params = isar.RadarParams(M, N, radar["band_width"], radar["fc"], radar["Tp"], radar["Fs"])
    out = {}
    for mode in (isar.AF_FUSED, isar.AF_SPLIT, isar.AF_SPLIT_PHASE, isar.AF_RESIDENT):
        with isar.Imager(params, max_frames=2) as im:
            im.set_autofocus_mode(mode)
            img = im.run_host(echo, vel, isar.Options())
            out[mode] = (img, [im.trace(k) for k in range(2)])
    img_f, tf = out[isar.AF_FUSED]
    for mode in (isar.AF_SPLIT, isar.AF_SPLIT_PHASE, isar.AF_RESIDENT):
        img_s, ts = out[mode]
        for k in range(2):
            assert ts[k].status & 2 == 0, f"mode {mode} frame {k}: resident wait bound hit"
            assert tf[k].unit2_num == ts[k].unit2_num
            assert np.array_equal(tf[k].phi, ts[k].phi), f"mode {mode} frame {k}: phases differ"
        assert np.array_equal(img_f, img_s), f"mode {mode}"


def test_resident_wait_bound_sets_status():
    """A resident-autofocus wait that gives up (bound forced negative from the host) marks the
    frame with status bit 2 and every workgroup still finishes; with the default bound no frame
    is marked and the phases equal the split form's."""
    cfgs = [synth.preset(1, seed=310 + k) for k in range(2)]
    frames = [synth.frame(c) for c in cfgs]
    radar = synth.radar_header(cfgs[0])
    M, N = cfgs[0].echo_num, cfgs[0].range_num
    echo = np.stack([f[0] for f in frames])
    vel = np.stack([f[1] for f in frames])
    print(f"[case] resident wait bound M={M} N={N}", flush=True)
    params = isar.RadarParams(M, N, radar["band_width"], radar["fc"], radar["Tp"], radar["Fs"])
    phis = {}
    for mode, polls in ((isar.AF_SPLIT, None), (isar.AF_RESIDENT, -1), (isar.AF_RESIDENT, None)):
        with isar.Imager(params, max_frames=2) as im:
            im.set_autofocus_mode(mode)
            if polls is not None:
                im.set_af_wait_polls(polls)
            im.run_host(echo, vel, isar.Options())
            tr = [im.trace(k) for k in range(2)]
        forced = polls is not None
        assert all(bool(t.status & 2) == forced for t in tr), [t.status for t in tr]
        phis[(mode, forced)] = [t.phi for t in tr]
    for k in range(2):
        assert np.array_equal(phis[(isar.AF_SPLIT, False)][k], phis[(isar.AF_RESIDENT, False)][k])


def test_batch_matches_single_frames():
    cfgs = [synth.preset(1, seed=100 + k) for k in range(3)]
    frames = [synth.frame(c) for c in cfgs]
    radar = synth.radar_header(cfgs[0])
    M, N = cfgs[0].echo_num, cfgs[0].range_num
    echo = np.stack([f[0] for f in frames])
    vel = np.stack([f[1] for f in frames])
    img_b, tb = _run_device(echo, vel, radar, M, N, isar.Options(), frames=3)
    for k in range(3):
        img_s, (ts,) = _run_device(echo[k], vel[k], radar, M, N, isar.Options())
        assert np.array_equal(img_b[k], img_s), f"frame {k} differs between batch and single runs"
        assert tb[k].cut_offset == ts.cut_offset and np.array_equal(tb[k].select_bin, ts.select_bin)


@pytest.mark.parametrize("cid,over", [(1, {}), (1, dict(echo_num=192)), (2, {})], ids=["c1", "c1_M192", "c2"])
def test_ra_update_launch_forms_agree(cid, over):
    """One alignment-update grid per merge level, or the rows of pair 0 apart from the other
    pairs (isar_set_ra_split), on a single frame: the same lags, phases and image bit for bit."""
    cfg, echo, vel, radar = _frame(cid, **over)
    M, N = cfg.echo_num, cfg.range_num
    print(f"[case] ra launch forms c{cid} M={M} N={N}", flush=True)
    params = isar.RadarParams(M, N, radar["band_width"], radar["fc"], radar["Tp"], radar["Fs"])
    out = []
    for split in (1, 2):
        with isar.Imager(params, max_frames=1) as im:
            im.set_ra_split(split)
            out.append((im.run_host(echo, vel, isar.Options()), im.trace(0)))
    (a, ta), (b, tb) = out
    assert np.array_equal(ta.ra_lag, tb.ra_lag) and np.array_equal(ta.phi, tb.phi)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("M,N", [(64, 1024), (128, 1024), (192, 1024), (256, 2048), (384, 1024), (512, 1024),
                                 (1024, 2048)])
def test_czt_lengths_and_tiles(M, N):
    """Every chirp-z length the context accepts (nextPow2(2M - 1), and 384 or 768 for 192 and 384
    pulses) against the oracle's chirp-z transform (mtrc.cu:125-192, any length gives the same
    transform), and both column-tile widths bit for bit; the lengths agree to FP32 rounding."""
    cfg, echo, vel, radar = _frame(1, echo_num=M, range_num=N)
    params = isar.RadarParams(M, N, radar["band_width"], radar["fc"], radar["Tp"], radar["Fs"])
    img_o, tr = oracle_run(echo, vel, radar, M, N)
    with isar.Imager(params, max_frames=1) as im:
        lengths = im.czt_lengths()
        p2 = 1 << (2 * M - 1).bit_length()
        assert lengths[0] == p2 and im.czt_length == p2
        assert lengths[1:] == ([384] if M == 192 else [768] if M == 384 else [])
        with pytest.raises(isar.ImagingError):
            im.set_czt_length(p2 * 2)
        out = {}
        for length in lengths:
            im.set_czt_length(length)
            for threads in (256, 512):
                print(f"[case] chirp-z M={M} N={N} length {length} threads {threads}", flush=True)
                im.set_czt_threads(threads)
                img = im.run_host(echo, vel, isar.Options())
                t = im.trace(0)
                rep = compare(img, t, img_o, tr, M)
                assert rep.ok(), f"length {length} threads {threads}: {rep}"
                out[length, threads] = img
        for length in lengths:
            assert np.array_equal(out[length, 256], out[length, 512]), f"length {length}: tile widths differ"
        ref = out[lengths[0], 256]
        for length in lengths[1:]:
            rel = np.linalg.norm(out[length, 256] - ref) / np.linalg.norm(ref)
            print(f"length {length} vs {lengths[0]}: rel L2 {rel:.1e}")
            assert rel < 1e-5


@pytest.mark.parametrize("M", [192, 256])
def test_autotune_keeps_a_valid_configuration(M):
    """isar_autotune over chunk rows, autofocus form, chirp-z length and column tile: the kept
    configuration is one of the candidates and images within the parity bars."""
    import torch
    K = 4
    cfgs = [synth.preset(1, echo_num=M, seed=900 + k) for k in range(K)]
    frames = [synth.frame(c) for c in cfgs]
    radar = synth.radar_header(cfgs[0])
    N = cfgs[0].range_num
    echo, vel = np.stack([f[0] for f in frames]), np.stack([f[1] for f in frames])
    params = isar.RadarParams(M, N, radar["band_width"], radar["fc"], radar["Tp"], radar["Fs"])
    dev = torch.device("cuda:0")
    d_echo = torch.from_numpy(echo.view(np.float32)).to(dev)
    d_vel = torch.from_numpy(vel).to(dev)
    d_img = torch.zeros((K, M, 512), dtype=torch.float32, device=dev)
    print(f"[case] autotune M={M} x {K} frames", flush=True)
    with isar.Imager(params, max_frames=K) as im:
        s = torch.cuda.current_stream().cuda_stream
        rows = im.autotune(d_echo.data_ptr(), d_vel.data_ptr(), d_img.data_ptr(), K, repeats=2, stream=s)
        print(f"chunk {rows}, autofocus {im.autofocus_mode}, chirp-z {im.czt_length} x {im.czt_threads}")
        assert rows in (2, 4, 8, 16)
        assert im.czt_length in im.czt_lengths() and im.czt_threads in (256, 512)
        img = im.run_host(echo, vel, isar.Options())
        for k in range(K):
            img_o, tr = oracle_run(echo[k], vel[k], radar, M, N)
            rep = compare(img[k], im.trace(k), img_o, tr, M)
            assert rep.ok(), f"frame {k}: {rep}"
