"""The frame loop (hmFrameLoopRun): host BGR frames in, stitched BGR frames out,
in order, each equal to the pano's own stitch of that frame (for a 4-channel
pano: the reference loop's BGR2BGRA -> process -> BGRA2BGR), checked against
the oracle; every ring depth; early stop and callback errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bgr_frames(rig, batch, seed):
    import hm_synth

    return [np.stack([hm_synth.make_frame(rig.src_h, rig.src_w, seed + 31 * i + 5 * b, 3) for b in range(batch)])
            for i in range(rig.n)]


def bgra(f):
    return np.concatenate([f, np.full(f.shape[:-1] + (1,), 255, np.uint8)], -1)


def run_loop(hm, pano, rig, n_frames, slots, batch=1, max_frames=-1, upload=0, download=0):
    frames = {k: bgr_frames(rig, batch, 1000 + 17 * k) for k in range(n_frames)}
    got = {}

    def source(k):
        return frames.get(k)

    def sink(k, bgr):
        got[k] = bgr.copy()

    st = hm.frame_loop(pano, [(rig.src_w, rig.src_h)] * rig.n, source, sink, slots=slots, max_frames=max_frames,
                       upload=upload, download=download)
    return frames, got, st


@pytest.mark.parametrize("slots", [1, 2, 3])
@pytest.mark.parametrize("levels", [0, 4])
def test_two_image_loop_matches_oracle(hm, torch_gpu, oracle, slots, levels):
    import hm_synth

    rig = hm_synth.small_rig(seed=151 + levels)
    pano = hm.StitchPano(1, levels, rig)
    frames, got, st = run_loop(hm, pano, rig, 7, slots)
    assert st["frames"] == 7 and sorted(got) == list(range(7))
    for k, f in frames.items():
        want = oracle.pano2(rig, bgra(f[0]), bgra(f[1]), levels)[0, ..., :3]
        np.testing.assert_array_equal(got[k], want, err_msg=f"frame {k}")
    w, h = pano.canvas_size()
    assert st["bytes_down_per_frame"] == w * h * 3
    assert st["bytes_up_per_frame"] == 2 * rig.src_w * rig.src_h * 3
    # the steady part: every frame after the first, within the call's wall time
    assert st["steady_frames"] == 6
    assert 0 < st["first_frame_seconds"] and 0 < st["steady_seconds"]
    assert st["first_frame_seconds"] + st["steady_seconds"] <= st["seconds"] + 1e-6


@pytest.mark.parametrize("upload,download", [(1, 0), (0, 1), (1, 1)])
def test_kernel_transfers(hm, torch_gpu, oracle, upload, download):
    """Copy kernels over mapped pinned memory instead of DMA: the same frames."""
    import hm_synth

    rig = hm_synth.small_rig(seed=155)
    pano = hm.StitchPano(1, 3, rig)
    frames, got, _ = run_loop(hm, pano, rig, 5, 3, upload=upload, download=download)
    for k, f in frames.items():
        want = oracle.pano2(rig, bgra(f[0]), bgra(f[1]), 3)[0, ..., :3]
        np.testing.assert_array_equal(got[k], want, err_msg=f"frame {k}")


def test_three_channel_pano_loop(hm, torch_gpu, oracle):
    """A uchar3 pano stitches the BGR frames directly (no expansion)."""
    import hm_synth

    rig = hm_synth.small_rig(seed=161)
    pano = hm.StitchPano(1, 3, rig, channels=3)
    frames, got, _ = run_loop(hm, pano, rig, 4, 2)
    for k, f in frames.items():
        np.testing.assert_array_equal(got[k], oracle.pano2(rig, f[0], f[1], 3)[0], err_msg=f"frame {k}")


def test_n_image_batched_loop(hm, torch_gpu, oracle):
    import hm_synth

    rig = hm_synth.small_rig(n=4, seed=171, src=(120, 56), canvas=(440, 64))
    pano = hm.StitchPanoN(2, 3, rig)
    frames, got, st = run_loop(hm, pano, rig, 5, 3, batch=2, max_frames=4)
    assert sorted(got) == [0, 1, 2, 3] and st["frames"] == 4  # max_frames stops before the source ends
    w, h = pano.canvas_size()
    for k in range(4):
        want = oracle.panoN(rig, [bgra(x) for x in frames[k]], 3)[..., :3].reshape(2 * h, w, 3)
        np.testing.assert_array_equal(got[k], want, err_msg=f"frame {k}")


def test_three_image_loop_keeps_the_canvas_between_frames(hm, torch_gpu, oracle):
    """The 3-image class leaves columns beyond a narrower seam untouched: like the
    reference's loop, one canvas carries them from frame to frame (black at first)."""
    import hm_synth

    rig = hm_synth.small_rig(n=3, seed=181, src=(120, 56), canvas=(330, 64))
    rig.seam = np.ascontiguousarray(rig.seam[:, : rig.seam.shape[1] - 29])
    pano = hm.StitchPano3(1, 2, rig)
    frames, got, _ = run_loop(hm, pano, rig, 3, 3)
    w, h = pano.canvas_size()
    prior = np.zeros((1, h, w, 4), np.uint8)
    for k in range(3):
        want = oracle.pano3(rig, [bgra(x) for x in frames[k]], 2, fused=True, canvas=prior)
        np.testing.assert_array_equal(got[k], want[0, ..., :3], err_msg=f"frame {k}")
        prior = want


def test_sink_stop_and_source_error(hm, torch_gpu):
    import hm_synth

    rig = hm_synth.small_rig(seed=191)
    pano = hm.StitchPano(1, 2, rig)
    seen = []

    def sink(k, bgr):
        seen.append(k)
        return False if k == 2 else None  # stop after frame 2

    st = hm.frame_loop(pano, [(rig.src_w, rig.src_h)] * 2, lambda k: bgr_frames(rig, 1, k), sink, slots=3)
    assert seen == [0, 1, 2] and st["frames"] == 3

    def bad_source(k):
        if k == 3:
            raise KeyError("no frame 3")
        return bgr_frames(rig, 1, k)

    with pytest.raises(KeyError):
        hm.frame_loop(pano, [(rig.src_w, rig.src_h)] * 2, bad_source, lambda k, b: None, slots=2)
    with pytest.raises(ValueError):
        hm.frame_loop(pano, [(rig.src_w, rig.src_h)] * 2, lambda k: [np.zeros((3, 3, 3), np.uint8)] * 2,
                      lambda k, b: None)
    # the pano still works afterwards
    frames, got, _ = run_loop(hm, pano, rig, 2, 2)
    assert sorted(got) == [0, 1]


def test_raw_video_command_line(hm, torch_gpu, oracle, tmp_path):
    """hm-cupano_amd/stitch_raw_video.py: a control directory and raw bgr24 streams in,
    the raw bgr24 panorama stream out."""
    import os
    import subprocess
    import sys

    import hm_synth
    from PIL import Image

    rig = hm_synth.small_rig(seed=201)
    mx, my = min(rig.pos_x), min(rig.pos_y)
    rig.pos_x = [p - mx for p in rig.pos_x]
    rig.pos_y = [p - my for p in rig.pos_y]
    d = str(tmp_path / "rig")
    os.makedirs(d)
    for i in range(2):
        Image.fromarray(rig.map_x[i]).save(os.path.join(d, f"mapping_{i:04d}_x.tif"), compression="tiff_lzw")
        Image.fromarray(rig.map_y[i]).save(os.path.join(d, f"mapping_{i:04d}_y.tif"), compression="tiff_lzw")
        Image.fromarray(np.zeros((2, 2, 3), np.uint8)).save(
            os.path.join(d, f"mapping_{i:04d}.tif"),
            tiffinfo={282: 1.0, 283: 1.0, 286: float(rig.pos_x[i]), 287: float(rig.pos_y[i])})
    Image.fromarray(np.where(rig.seam == 1, 0, 255).astype(np.uint8)).save(os.path.join(d, "seam_file.png"))
    frames = [bgr_frames(rig, 1, 300 + 9 * k) for k in range(5)]
    paths = [str(tmp_path / f"cam{i}.bgr") for i in range(2)]
    for i, p in enumerate(paths):
        with open(p, "wb") as f:
            for fr in frames:
                f.write(fr[i].tobytes())
    out = str(tmp_path / "pano.bgr")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "hm-cupano_amd", "stitch_raw_video.py"), "--control", d,
                        "--input", paths[0], "--input", paths[1], "--input-size", f"{rig.src_w}x{rig.src_h}",
                        "--levels", "4", "--output", out], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    assert "Processed frames: 5" in r.stderr
    w, h = (int(v) for v in r.stderr.split("Canvas: ")[1].split()[0].split("x"))
    data = np.fromfile(out, np.uint8).reshape(5, h, w, 3)
    for k, fr in enumerate(frames):
        want = oracle.pano2(rig, bgra(fr[0]), bgra(fr[1]), 4)[0, ..., :3]
        np.testing.assert_array_equal(data[k], want, err_msg=f"frame {k}")
