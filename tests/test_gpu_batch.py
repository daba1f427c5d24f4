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


def oracle_out(oracle, kind, rig, frames, levels, fused, prior):
    if kind == "2":
        return oracle.pano2(rig, frames[0], frames[1], levels)
    if kind == "N":
        return oracle.panoN(rig, frames, levels)
    return oracle.pano3(rig, frames, levels, fused=fused, canvas=prior)


def build(hm, torch, calls, seed):
    bufs = []
    for k, (name, pano, rig, levels, kind, fused, b) in enumerate(calls):
        fr = frames_for(rig, b, seed + 10 * k)
        w, h = pano.canvas_size()
        prior = np.full((b, h, w, 4), 7 + k, np.uint8)
        bufs.append((fr, to_dev(torch, fr), torch.from_numpy(prior).cuda(), prior))
    torch.cuda.synchronize()
    batch = hm.Batch([(c[1], bufs[k][1], bufs[k][2], c[5]) for k, c in enumerate(calls)])
    return batch, bufs


def check(oracle, calls, bufs, msg):
    for k, (name, pano, rig, levels, kind, fused, b) in enumerate(calls):
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
    r3c = hm_synth.small_rig(seed=52)
    p3c = hm.StitchPano(1, 3, r3c, channels=3)
    with pytest.raises(hm.HmError):  # 3-channel pipelines are not carried
        hm.Batch([(p3c, to_dev(torch, frames_for(r3c, 1, 2, channels=3)), p3c.alloc_canvas())])
    r5 = hm_synth.small_rig(n=5, seed=53, src=(120, 56), canvas=(510, 64))
    p5 = hm.StitchPanoN(1, 3, r5)
    with pytest.raises(hm.HmError):  # five images: not carried
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
