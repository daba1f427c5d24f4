"""The batched pass over mixed configurations (hmBatch*): pano calls of the 2-,
3- and N-image classes with different image counts, canvas sizes, level
counts and batch sizes, run as one persistent launch per stage.  Every frame
must be bit-identical to the same frame stitched alone (the oracle)."""
import numpy as np
import pytest

from conftest import frames_for

pytestmark = pytest.mark.gpu


def to_dev(torch, frames):
    return [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]


def make_calls(hm):
    """(name, pano, rig, levels, kind, fused, batch) for a mix the pass carries."""
    import hm_synth

    out = []
    r = hm_synth.small_rig(seed=31)
    out.append(("pano2_L6_b2", hm.StitchPano(2, 6, r), r, 6, "2", True, 2))
    r = hm_synth.small_rig(seed=32, src=(200, 80), canvas=(520, 96))
    out.append(("pano2_L3", hm.StitchPano(1, 3, r), r, 3, "2", True, 1))
    r = hm_synth.small_rig(seed=33)
    out.append(("pano2_L0", hm.StitchPano(1, 0, r), r, 0, "2", True, 1))
    r = hm_synth.small_rig(n=3, seed=34, src=(120, 56), canvas=(330, 64))
    out.append(("panoN3_L3", hm.StitchPanoN(1, 3, r), r, 3, "N", True, 1))
    r = hm_synth.small_rig(n=4, seed=35, src=(120, 56), canvas=(420, 64))
    out.append(("panoN4_L4_b2", hm.StitchPanoN(2, 4, r), r, 4, "N", True, 2))
    r = hm_synth.small_rig(n=4, seed=36, src=(120, 56), canvas=(420, 64))
    out.append(("panoN4_L0", hm.StitchPanoN(1, 0, r), r, 0, "N", True, 1))
    r = hm_synth.small_rig(n=3, seed=37, src=(120, 56), canvas=(330, 64))
    out.append(("pano3_L4_fused", hm.StitchPano3(1, 4, r), r, 4, "3", True, 1))
    r = hm_synth.small_rig(n=3, seed=38, src=(120, 56), canvas=(330, 64))
    out.append(("pano3_L2_unfused", hm.StitchPano3(1, 2, r), r, 2, "3", False, 1))
    return out


def make_calls_wide(hm):
    """A mix of 3-channel pipelines of every class and of N-image rigs with 2 and
    5 to 8 images: (name, pano, rig, levels, kind, fused, batch, channels)."""
    import hm_synth

    def n_rig(n, seed):
        return hm_synth.small_rig(n=n, seed=seed, src=(120, 56), canvas=(90 * n + 60, 64))

    out = []
    r = hm_synth.small_rig(seed=61)
    out.append(("pano2_c3_L5", hm.StitchPano(1, 5, r, channels=3), r, 5, "2", True, 1, 3))
    r = hm_synth.small_rig(seed=62)
    out.append(("pano2_c3_L0_b2", hm.StitchPano(2, 0, r, channels=3), r, 0, "2", True, 2, 3))
    r = n_rig(3, 63)
    out.append(("panoN3_c3_L3", hm.StitchPanoN(1, 3, r, channels=3), r, 3, "N", True, 1, 3))
    r = n_rig(5, 64)
    out.append(("panoN5_c3_L4", hm.StitchPanoN(1, 4, r, channels=3), r, 4, "N", True, 1, 3))
    r = hm_synth.small_rig(n=3, seed=65, src=(120, 56), canvas=(330, 64))
    out.append(("pano3_c3_L2_fused", hm.StitchPano3(1, 2, r, channels=3), r, 2, "3", True, 1, 3))
    r = hm_synth.small_rig(n=3, seed=66, src=(120, 56), canvas=(330, 64))
    out.append(("pano3_c3_L3_unfused", hm.StitchPano3(1, 3, r, channels=3), r, 3, "3", False, 1, 3))
    r = n_rig(2, 67)
    out.append(("panoN2_L3", hm.StitchPanoN(1, 3, r), r, 3, "N", True, 1, 4))
    r = n_rig(5, 68)
    out.append(("panoN5_L0", hm.StitchPanoN(1, 0, r), r, 0, "N", True, 1, 4))
    r = n_rig(6, 69)
    out.append(("panoN6_L4_b2", hm.StitchPanoN(2, 4, r), r, 4, "N", True, 2, 4))
    r = n_rig(7, 70)
    out.append(("panoN7_L2", hm.StitchPanoN(1, 2, r), r, 2, "N", True, 1, 4))
    r = n_rig(8, 71)
    out.append(("panoN8_L6", hm.StitchPanoN(1, 6, r), r, 6, "N", True, 1, 4))
    r = n_rig(8, 72)
    out.append(("panoN8_c3_L5", hm.StitchPanoN(1, 5, r, channels=3), r, 5, "N", True, 1, 3))
    return out


def oracle_out(oracle, kind, rig, frames, levels, fused, prior):
    if kind == "2":
        return oracle.pano2(rig, frames[0], frames[1], levels)
    if kind == "N":
        return oracle.panoN(rig, frames, levels)
    return oracle.pano3(rig, frames, levels, fused=fused, canvas=prior)


def build(hm, torch, calls, seed):
    bufs = []
    for k, (name, pano, rig, levels, kind, fused, b, *ch) in enumerate(calls):
        c = ch[0] if ch else 4
        fr = frames_for(rig, b, seed + 10 * k, c)
        w, h = pano.canvas_size()
        prior = np.full((b, h, w, c), 7 + k, np.uint8)
        bufs.append((fr, to_dev(torch, fr), torch.from_numpy(prior).cuda(), prior))
    torch.cuda.synchronize()
    batch = hm.Batch([(c[1], bufs[k][1], bufs[k][2], c[5]) for k, c in enumerate(calls)])
    return batch, bufs


def check(oracle, calls, bufs, msg):
    for k, (name, pano, rig, levels, kind, fused, b, *_) in enumerate(calls):
        fr, _, canvas, prior = bufs[k]
        want = oracle_out(oracle, kind, rig, fr, levels, fused, prior)
        np.testing.assert_array_equal(canvas.cpu().numpy(), want, err_msg=f"{name} {msg}")


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_mixed_matches_alone(hm, torch_gpu, oracle, use_graph):
    torch = torch_gpu
    calls = make_calls(hm)
    batch, bufs = build(hm, torch, calls, 100)
    info = batch.info()
    assert info["launches"] < info["recorded_launches"]
    assert info["stages"] <= info["launches"]
    assert batch.run(use_graph=use_graph) == 0
    torch.cuda.synchronize()
    check(oracle, calls, bufs, f"graph={use_graph}")


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_three_channels_and_up_to_eight_images(hm, torch_gpu, oracle, use_graph):
    """3-channel pipelines of every class and N-image rigs of 2 and 5 to 8 images in one pass."""
    torch = torch_gpu
    calls = make_calls_wide(hm)
    batch, bufs = build(hm, torch, calls, 500)
    info = batch.info()
    assert info["launches"] < info["recorded_launches"]
    assert batch.run(use_graph=use_graph) == 0
    torch.cuda.synchronize()
    check(oracle, calls, bufs, f"graph={use_graph}")


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_two_image_up_pass_shapes(hm, torch_gpu, oracle, use_graph):
    """2-image rigs under every up-pass shape the pass carries: the burst tail (base 1..3)
    and the per-level tail form (more levels than the burst holds), with levels 1 and 0 in
    one launch or two, 3 and 4 channels, levels 2..8."""
    import hm_synth

    torch = torch_gpu
    cases = [  # (levels, channels, tail_base, fuse_up)
        (6, 4, 3, 1), (6, 4, 2, 0), (6, 3, 1, 1), (4, 4, 2, 1), (4, 3, 0, 0), (2, 4, 2, 1), (8, 4, 1, 1),
        (8, 3, 2, 0),
    ]
    calls = []
    for k, (levels, ch, tail, fu) in enumerate(cases):
        r = hm_synth.small_rig(seed=80 + k, src=(200, 96), canvas=(520, 112))
        p = hm.StitchPano(1, levels, r, channels=ch)
        p.set_tuning(tail_base=tail, fuse_up=fu)
        calls.append((f"pano2_L{levels}_c{ch}_tail{tail}_fu{fu}", p, r, levels, "2", True, 1, ch))
    batch, bufs = build(hm, torch, calls, 900)
    assert batch.run(use_graph=use_graph) == 0
    torch.cuda.synchronize()
    check(oracle, calls, bufs, f"graph={use_graph}")
    for (levels, ch, tail, fu), (name, p, *_r) in zip(cases, calls):  # the recording leaves the tuning alone
        t = p.get_tuning()
        assert (t["tail_base"], t["fuse_up"]) == (tail, fu), name


def test_batch_stage_split_over_several_launches(hm, torch_gpu, oracle):
    """More calls of one body than a launch carries: 18 2-image rigs (more than the 16 items of
    a launch) and 6 8-image rigs (their argument blocks overflow a launch's 3.5 KB), so each
    stage runs several launches of the same body; every frame equals the oracle's."""
    import hm_synth

    torch = torch_gpu
    calls = []
    for k in range(18):
        r = hm_synth.small_rig(seed=200 + k)
        calls.append((f"pano2_{k}", hm.StitchPano(1, 2 + k % 3, r), r, 2 + k % 3, "2", True, 1))
    for k in range(6):
        r = hm_synth.small_rig(n=8, seed=300 + k, src=(120, 56), canvas=(780, 64))
        calls.append((f"panoN8_{k}", hm.StitchPanoN(1, 3, r), r, 3, "N", True, 1))
    batch, bufs = build(hm, torch, calls, 1200)
    info = batch.info()
    assert info["launches"] < info["recorded_launches"]
    for use_graph in (False, True):
        for _, _, c, prior in bufs:
            c.copy_(torch.from_numpy(prior).cuda())
        torch.cuda.synchronize()
        assert batch.run(use_graph=use_graph) == 0
        torch.cuda.synchronize()
        check(oracle, calls, bufs, f"graph={use_graph}")


def test_batch_more_than_four_images_per_tile(hm, torch_gpu, oracle):
    """Six images over the whole canvas (identity maps, all at (0, 0)): every tile
    meets all six, so the batched down pass runs several slot chunks per tile and
    the up pass its per-pixel path; beside it an 8-image rig with a 2-image one."""
    import hm_synth
    from conftest import IdRig

    torch = torch_gpu
    w, h, n = 160, 72, 6
    ys, xs = np.mgrid[0:h, 0:w]
    seam = np.clip(((xs + 7 * np.sin(ys / 5.0)) * n) // w, 0, n - 1).astype(np.uint8)
    rig = IdRig([(w, h)] * n, [(0, 0)] * n, seam)
    rng = np.random.default_rng(81)
    imgs = []
    for i in range(n):
        im = rng.integers(0, 256, (1, h, w, 4), dtype=np.uint8)
        im[..., 3] = 255
        im[0, 10 * i: 10 * i + 6, 20:40, 3] = 0  # a transparent hole per image
        imgs.append(im)
    pano6 = hm.StitchPanoN(1, 4, rig)
    r8 = hm_synth.small_rig(n=8, seed=82, src=(120, 56), canvas=(780, 64))
    pano8 = hm.StitchPanoN(1, 3, r8)
    fr8 = frames_for(r8, 1, 83)
    r2 = hm_synth.small_rig(seed=84)
    pano2 = hm.StitchPano(1, 4, r2)
    fr2 = frames_for(r2, 1, 85)
    cv = [p.alloc_canvas() for p in (pano6, pano8, pano2)]
    batch = hm.Batch([(pano6, to_dev(torch, imgs), cv[0]), (pano8, to_dev(torch, fr8), cv[1]),
                      (pano2, to_dev(torch, fr2), cv[2])])
    assert batch.run(use_graph=True) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(cv[0].cpu().numpy(), oracle.panoN(rig, imgs, 4))
    np.testing.assert_array_equal(cv[1].cpu().numpy(), oracle.panoN(r8, fr8, 3))
    np.testing.assert_array_equal(cv[2].cpu().numpy(), oracle.pano2(r2, fr2[0], fr2[1], 4))


@pytest.mark.parametrize("grid", [1, 7, 64])
def test_batch_any_grid(hm, torch_gpu, oracle, grid):
    """The work-list stride: a grid of one workgroup walks every tile itself."""
    torch = torch_gpu
    calls = make_calls(hm)[:5]
    batch, bufs = build(hm, torch, calls, 200)
    batch.set_grid(grid)
    assert batch.info()["grid"] == grid
    assert batch.run() == 0
    torch.cuda.synchronize()
    check(oracle, calls, bufs, f"grid={grid}")


def test_batch_graph_replays_with_new_content(hm, torch_gpu, oracle):
    torch = torch_gpu
    calls = make_calls(hm)
    batch, bufs = build(hm, torch, calls, 300)
    for rep in range(3):
        for k, (name, pano, rig, levels, kind, fused, b) in enumerate(calls):
            fr = frames_for(rig, b, 1000 + 37 * rep + k)
            for d, f in zip(bufs[k][1], fr):
                d.copy_(torch.from_numpy(f).cuda())  # eager work on torch's stream
            bufs[k] = (fr,) + bufs[k][1:]
        torch.cuda.synchronize()
        assert batch.run(use_graph=True) == 0
        torch.cuda.synchronize()
        scratch = torch.ones(1 << 20, device="cuda").sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert float(scratch) == float(1 << 20)
        check(oracle, calls, bufs, f"replay {rep}")


def test_batch_leaves_tuning_and_plain_path_alone(hm, torch_gpu, oracle):
    """Recording forces the carried launch shapes for itself only."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=41)
    pano = hm.StitchPano(1, 5, rig)
    pano.set_tuning(remap_px=8, tail_base=3, overlap=1, remap_xcd=0, remap_rows=4, canvas_grid=256, fuse_down=1)
    before = pano.get_tuning()
    fr = frames_for(rig, 1, 5)
    dev = to_dev(torch, fr)
    canvas = pano.alloc_canvas()
    batch = hm.Batch([(pano, dev, canvas)])
    assert pano.get_tuning() == before
    assert batch.run() == 0
    torch.cuda.synchronize()
    want = oracle.pano2(rig, fr[0], fr[1], 5)
    np.testing.assert_array_equal(canvas.cpu().numpy(), want)
    canvas.fill_(0)
    assert pano.process(dev[0], dev[1], canvas) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(canvas.cpu().numpy(), want)


def test_batch_errors(hm, torch_gpu):
    import hm_synth

    torch = torch_gpu
    r = hm_synth.small_rig(seed=51)
    pano = hm.StitchPano(1, 3, r)
    dev = to_dev(torch, frames_for(r, 1, 1))
    canvas = pano.alloc_canvas()
    with pytest.raises(hm.HmError):  # one object twice in a pass
        hm.Batch([(pano, dev, canvas), (pano, dev, canvas)])
    r5 = hm_synth.small_rig(n=5, seed=53, src=(120, 56), canvas=(510, 64))
    p5 = hm.StitchPanoN(1, 3, r5, debug=True)
    with pytest.raises(hm.HmError):  # debug mode (every level kept) is not carried
        hm.Batch([(p5, to_dev(torch, frames_for(r5, 1, 3)), p5.alloc_canvas())])
    # a failed create leaves the panos usable on their own
    assert pano.process(dev[0], dev[1], canvas) == 0


def test_batch_autotune_keeps_output(hm, torch_gpu, oracle):
    torch = torch_gpu
    calls = make_calls(hm)
    batch, bufs = build(hm, torch, calls, 400)
    grid, ms = batch.autotune(passes=2)
    assert ms > 0.0
    assert batch.info()["grid"] >= 1
    assert batch.run(use_graph=True) == 0
    torch.cuda.synchronize()
    check(oracle, calls, bufs, f"autotuned grid={grid}")
