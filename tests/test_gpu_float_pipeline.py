"""The float pipelines of the pano classes (CudaStitchPano / CudaStitchPanoN /
CudaStitchPano3 <float4|float3, float4|float3>) on the GPU against the oracle:
float inputs with fractional values, a float canvas, no truncation.  The
results are compared exactly; the reference's own float pano tests
(src/pano/cudaPano_test.cpp:141-204, cudaPanoN_test.cpp:192-299,
cudaPano3_test.cpp:133-196) are restated at the end."""
import numpy as np
import pytest

from conftest import IdRig, frames_for, patterned_u8

pytestmark = pytest.mark.gpu


def float_frames(rig, batch, seed, channels=4):
    """Frames as float with fractional parts (the u8 frame plus a seeded fraction)."""
    rng = np.random.default_rng(seed)
    out = []
    for f in frames_for(rig, batch, seed, channels):
        v = f.astype(np.float32) + rng.uniform(0, 1, f.shape).astype(np.float32)
        if channels == 4:
            v[..., 3] = 255.0
        out.append(v)
    return out


def to_dev(torch, frames):
    return [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]


def run(torch, pano, kind, frames, canvas_fill=7.0, fused=True):
    w, h = pano.canvas_size()
    b, c = frames[0].shape[0], frames[0].shape[3]
    canvas = torch.full((b, h, w, c), canvas_fill, dtype=torch.float32, device="cuda")
    ins = to_dev(torch, frames)
    if kind == "2":
        rc = pano.process(ins[0], ins[1], canvas)
    elif kind == "3":
        rc = pano.process(ins[0], ins[1], ins[2], canvas, fused=fused)
    else:
        rc = pano.process(ins, canvas)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return canvas.cpu().numpy()


@pytest.mark.parametrize("levels", [0, 1, 3, 6])
@pytest.mark.parametrize("minimize", [True, False])
@pytest.mark.parametrize("channels", [4, 3])
def test_pano2_float_pipeline_matches_oracle(hm, torch_gpu, oracle, levels, minimize, channels):
    import hm_synth

    rig = hm_synth.small_rig(seed=11)
    frames = float_frames(rig, 1, 3, channels)
    pano = hm.StitchPano(1, levels, rig, minimize_blend=minimize, channels=channels, float_pipeline=True)
    assert pano.status == 0, pano.status
    got = run(torch_gpu, pano, "2", frames)
    want = oracle.pano2(rig, frames[0], frames[1], levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)


def test_pano2_float_pipeline_batch2_and_next_frame(hm, torch_gpu, oracle):
    import hm_synth

    rig = hm_synth.small_rig(seed=8)
    pano = hm.StitchPano(2, 4, rig, float_pipeline=True)
    for seed in (1, 2):  # the second frame must not see the first
        frames = float_frames(rig, 2, seed)
        got = run(torch_gpu, pano, "2", frames)
        np.testing.assert_array_equal(got, oracle.pano2(rig, frames[0], frames[1], 4))


@pytest.mark.parametrize("n", [3, 5])
@pytest.mark.parametrize("levels", [0, 3, 6])
@pytest.mark.parametrize("minimize", [True, False])
def test_panoN_float_pipeline_matches_oracle(hm, torch_gpu, oracle, n, levels, minimize):
    import hm_synth

    rig = hm_synth.small_rig(n=n, seed=30 + n, src=(120, 56), canvas=(90 * n + 60, 64))
    frames = float_frames(rig, 1, 5 + n)
    pano = hm.StitchPanoN(1, levels, rig, minimize_blend=minimize, float_pipeline=True)
    assert pano.status == 0, pano.status
    got = run(torch_gpu, pano, "N", frames)
    np.testing.assert_array_equal(got, oracle.panoN(rig, frames, levels, minimize=minimize))


@pytest.mark.parametrize("levels", [0, 2, 4])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("channels", [4, 3])
def test_pano3_float_pipeline_matches_oracle(hm, torch_gpu, oracle, levels, fused, channels):
    import hm_synth

    rig = hm_synth.small_rig(n=3, seed=50 + levels, src=(120, 56), canvas=(330, 64))
    frames = float_frames(rig, 1, 4 + levels, channels)
    pano = hm.StitchPano3(1, levels, rig, channels=channels, float_pipeline=True)
    assert pano.status == 0, pano.status
    got = run(torch_gpu, pano, "3", frames, fused=fused)
    prior = np.full(got.shape, 7.0, np.float32)
    np.testing.assert_array_equal(got, oracle.pano3(rig, frames, levels, fused=fused, canvas=prior))


def test_float_pipeline_is_refused_where_only_bytes_travel(hm, torch_gpu):
    """The batched pass and the BGR frame loop carry the uint8 pipeline only."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=11)
    pano = hm.StitchPano(1, 3, rig, float_pipeline=True)
    frames = to_dev(torch, float_frames(rig, 1, 1))
    with pytest.raises(hm.HmError):
        hm.Batch([(pano, frames, pano.alloc_canvas())])
    assert pano.alloc_canvas().dtype == torch.float32
    # a uint8 surface is not a float canvas: the pitch does not hold the float row
    bad = torch.zeros((1, 8, 8, 4), dtype=torch.uint8, device="cuda")
    assert pano.process(frames[0], frames[1], bad) != 0


# ---- the reference's float pano tests ----
def pattern_f4(w, h, idx):
    return patterned_u8(w, h, idx).astype(np.float32)


def test_pano2_float_minimize_matches_full(hm, torch_gpu, oracle):
    """src/pano/cudaPano_test.cpp:141-204: minimized = full within 1e-4, a narrower level 0."""
    w, h, x2 = 384, 64, 192
    cw = w + x2
    seam = np.zeros((h, cw), np.uint8)
    seam[:, : cw // 2] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (x2, 0)], seam)
    imgs = [pattern_f4(w, h, 0), pattern_f4(w, h, 1)]
    full = hm.StitchPano(1, 4, rig, minimize_blend=False, float_pipeline=True)
    mini = hm.StitchPano(1, 4, rig, minimize_blend=True, float_pipeline=True)
    got_full = run(torch_gpu, full, "2", imgs)
    got_mini = run(torch_gpu, mini, "2", imgs)
    assert np.abs(got_full - got_mini).max() <= 1e-4
    assert full.levels()[0] == (cw, h) and mini.levels()[0][0] < cw
    np.testing.assert_array_equal(got_mini, oracle.pano2(rig, imgs[0], imgs[1], 4, minimize=True))


PANO_N_CASES = {
    # src/pano/cudaPanoN_test.cpp:192-299: (image w, h, positions, canvas w, columns of label 1, images)
    "two_way_overlap": (384, 384, [(0, 0), (0, 0)], 384, (192, 384), 2),
    "no_overlap": (192, 384, [(0, 0), (192, 0)], 384, (192, 384), 2),
    "gap": (128, 384, [(0, 0), (256, 0)], 384, (192, 384), 2),
    "multi_seam_intersection": (384, 384, [(0, 0)] * 4, 384, None, 4),
}


@pytest.mark.parametrize("case", sorted(PANO_N_CASES))
def test_panoN_float_reference_cases(hm, torch_gpu, oracle, case):
    w, h, pos, cw, band, n = PANO_N_CASES[case]
    seam = np.zeros((h, cw), np.uint8)
    if band:
        seam[:, band[0]: band[1]] = 1
    else:  # four quadrant labels in a 64 x 64 square at the centre
        cx, cy = w // 2, h // 2
        for y in range(cy - 32, cy + 32):
            for x in range(cx - 32, cx + 32):
                seam[y, x] = (1 if y >= cy else 0) * 2 + (1 if x >= cx else 0)
    rig = IdRig([(w, h)] * n, pos, seam)
    imgs = [pattern_f4(w, h, i) for i in range(n)]
    full = run(torch_gpu, hm.StitchPanoN(1, 4, rig, minimize_blend=False, float_pipeline=True), "N", imgs)
    mini = run(torch_gpu, hm.StitchPanoN(1, 4, rig, minimize_blend=True, float_pipeline=True), "N", imgs)
    assert np.abs(full - mini).max() <= 1e-4
    if case == "gap":  # the uncovered columns stay transparent
        assert (mini[0, :, 128:256, 3] == 0).all()
    np.testing.assert_array_equal(mini, oracle.panoN(rig, imgs, 4, minimize=True))


def test_pano3_float_one_hot_selects_middle_image(hm, torch_gpu):
    """src/pano/cudaPano3_test.cpp:133-196."""
    rig = IdRig([(1, 1)] * 3, [(0, 0)] * 3, np.uint8([[1]]))
    imgs = [np.float32([[[[v, v, v, 255.0]]]]) for v in (30.0, 60.0, 90.0)]
    got = run(torch_gpu, hm.StitchPano3(1, 1, rig, float_pipeline=True), "3", imgs)
    np.testing.assert_allclose(got[0, 0, 0], [60.0, 60.0, 60.0, 255.0], atol=1e-3)
