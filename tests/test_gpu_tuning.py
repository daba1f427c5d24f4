"""Launch-shape knobs and the autotuner: every setting must give the same,
bit-exact output (the oracle)."""
import itertools

import numpy as np
import pytest

from conftest import frames_for

pytestmark = pytest.mark.gpu


def to_dev(torch, frames):
    return [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]


@pytest.mark.parametrize("levels", [6, 0])
def test_pano2_every_tuning_is_bit_exact(hm, torch_gpu, oracle, levels):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=61)
    frames = frames_for(rig, 1, 5)
    want = oracle.pano2(rig, frames[0], frames[1], levels)
    pano = hm.StitchPano(1, levels, rig)
    dev = to_dev(torch, frames)
    shapes = itertools.product((2, 4, 8), (0, 1), (1, 2, 4), (0, 1))  # remap pixels, XCD order, rows, fused down
    # tail, overlap, canvas grid, fused down, levels 1 and 0 fused
    rest = itertools.product((0, 1, 2, 3), (0, 1, 2, 3), (0, 256, 2048), (0, 1), (0, 1))
    settings = [(px, 2, 1, xc, ry, 0, fd, 1) for px, xc, ry, fd in shapes] + \
        [(4, tb, ov, 1, 2, gd, fd, fu) for tb, ov, gd, fd, fu in rest]
    for px, tb, ov, xc, ry, gd, fd, fu in settings:
        pano.set_tuning(px, tb, ov, xc, ry, gd, fd, fu)
        canvas = torch.full(want.shape, 33, dtype=torch.uint8, device="cuda")
        assert pano.process(dev[0], dev[1], canvas) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(canvas.cpu().numpy(), want,
                                      err_msg=f"px={px} tail={tb} overlap={ov} xcd={xc} rows={ry} grid={gd} "
                                              f"fuse={fd} fuse_up={fu}")


def test_multi_image_overlap_knob_is_bit_exact(hm, torch_gpu, oracle):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(n=3, seed=62, src=(120, 56), canvas=(330, 64))
    frames = frames_for(rig, 1, 6)
    dev = to_dev(torch, frames)
    want_n = oracle.panoN(rig, frames, 4)
    pn = hm.StitchPanoN(1, 4, rig)
    rig3 = hm_synth.small_rig(n=3, seed=63, src=(120, 56), canvas=(330, 64))
    rig3.seam = np.ascontiguousarray(rig3.seam[:, :-29])  # blend narrower than the canvas: a side launch
    fr3 = frames_for(rig3, 1, 7)
    d3 = to_dev(torch, fr3)
    p3 = hm.StitchPano3(1, 3, rig3)
    for ov in (0, 1):
        pn.set_tuning(4, 2, ov)
        c = pn.alloc_canvas()
        assert pn.process(dev, c) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(c.cpu().numpy(), want_n, err_msg=f"panoN overlap={ov}")
        p3.set_tuning(4, 2, ov)
        w, h = p3.canvas_size()
        prior = np.full((1, h, w, 4), 5, np.uint8)
        c3 = torch.from_numpy(prior).cuda()
        torch.cuda.synchronize()
        assert p3.process(d3[0], d3[1], d3[2], c3, fused=False) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(c3.cpu().numpy(), oracle.pano3(rig3, fr3, 3, fused=False, canvas=prior),
                                      err_msg=f"pano3 overlap={ov}")


def test_autotune_picks_a_setting_and_stays_exact(hm, torch_gpu, oracle):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=64)
    frames = frames_for(rig, 1, 8)
    pano = hm.StitchPano(1, 5, rig)
    dev = to_dev(torch, frames)
    canvas = pano.alloc_canvas()
    best, ms = pano.autotune(dev, canvas, frames=3)
    assert best["remap_px"] in (2, 4, 8) and 0 <= best["tail_base"] <= 3 and best["overlap"] in (0, 1, 2, 3)
    assert ms > 0 and pano.get_tuning() == best
    canvas.fill_(0)
    assert pano.process(dev[0], dev[1], canvas) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.pano2(rig, frames[0], frames[1], 5))


def test_invalid_tuning_is_refused(hm, torch_gpu):
    import hm_synth

    pano = hm.StitchPano(1, 4, hm_synth.small_rig(seed=65))
    with pytest.raises(hm.HmError):
        pano.set_tuning(3, 2, 1)
    with pytest.raises(hm.HmError):
        pano.set_tuning(4, 7, 1)
    assert pano.get_tuning() == {"remap_px": 4, "tail_base": 2, "overlap": 1, "remap_xcd": 0, "remap_rows": 1,
                                 "canvas_grid": 0, "fuse_down": 0, "fuse_up": 1}
    with pytest.raises(hm.HmError):
        pano.set_tuning(4, 2, 1, 0, 1, 0, 0, 2)


@pytest.mark.parametrize("n", [3, 4, 5])
def test_multi_image_tail_and_fused_levels_are_bit_exact(hm, torch_gpu, oracle, n):
    """Three to five images at levels 6: the up pass's small levels in one burst launch from
    tail_base (three and four images) or one launch each (five), with and without levels 1
    and 0 fused, all equal to the oracle; the 3-image class's tail likewise."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(n=n, seed=90 + n, src=(120, 56), canvas=(90 * n + 60, 64))
    frames = frames_for(rig, 1, 11 + n)
    dev = to_dev(torch, frames)
    want = oracle.panoN(rig, frames, 6)
    pano = hm.StitchPanoN(1, 6, rig)
    for tb, fu in itertools.product((0, 1, 2, 3), (0, 1)):
        pano.set_tuning(4, tb, 1, 0, 1, 0, 0, fu)
        pano.set_timing(True)
        c = pano.alloc_canvas()
        assert pano.process(dev, c) == 0
        torch.cuda.synchronize()
        labels = set(pano.timing_report())
        assert ("up_tail" in labels) == (n <= 4 and tb > 0), (tb, fu, labels)
        np.testing.assert_array_equal(c.cpu().numpy(), want, err_msg=f"n={n} tail_base={tb} fuse_up={fu}")
    if n == 3:
        p3 = hm.StitchPano3(1, 6, rig)
        w3, h3 = p3.canvas_size()
        want3 = oracle.pano3(rig, frames, 6, fused=True, canvas=np.zeros((1, h3, w3, 4), np.uint8))
        for tb in (0, 2, 3):
            p3.set_tuning(4, tb, 1)
            c = torch.zeros(want3.shape, dtype=torch.uint8, device="cuda")
            assert p3.process(dev[0], dev[1], dev[2], c, fused=True) == 0
            torch.cuda.synchronize()
            np.testing.assert_array_equal(c.cpu().numpy(), want3, err_msg=f"pano3 tail_base={tb}")
