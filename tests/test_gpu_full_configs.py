"""The benchmark workloads at their full sizes, on the GPU, against the CPU oracle.

  * 2 x 5312x2988 -> 8500x3300 at levels 6 under the launch shapes the
    autotuner picks (and the other persistent / fused shapes);
  * 8 x 3840x2160 -> 14000x2600 at levels 6;
  * the 16-frame mixed batch (6 x 2-image 8500x3300 L=6, 6 x 3-image
    6000x2000 L=4, 4 x 4-image 4000x1500 L=0) as one hmBatch pass, every
    frame against the oracle;
  * the 32-rig served stream (5 configs, 2 to 8 images, levels 0/4/6) through
    hmServer: every frame equal to the same frame stitched alone, and one
    frame per config against the oracle.

Everything is bit-exact.  Rigs, frames and oracle outputs are generated once
per module: rigs of one config share their geometry and input sets, so the
oracle runs once per (config, input set)."""
import gc

import numpy as np
import pytest

from conftest import frames_for

pytestmark = pytest.mark.gpu

_rigs = {}
_frames = {}
_want = {}


def served_rig(name):
    import hm_synth

    if name not in _rigs:
        _rigs[name] = hm_synth.served_rig(name)
    return _rigs[name]


def served_frames(name, k):
    import hm_synth

    if (name, k) not in _frames:
        _frames[name, k] = hm_synth.served_frames(served_rig(name), k)
    return _frames[name, k]


def oracle_of(oracle, name, k, prior=None):
    """The oracle's frame k of a served config (3-image class: over `prior`)."""
    import hm_synth

    key = (name, k, None if prior is None else int(prior.flat[0]))
    if key not in _want:
        kind, levels, _, _ = hm_synth.served_config(name)
        rig, fr = served_rig(name), served_frames(name, k)
        print(f"oracle: {name} input set {k}", flush=True)
        if kind == "2":
            _want[key] = oracle.pano2(rig, fr[0], fr[1], levels)
        elif kind == "3":
            _want[key] = oracle.pano3(rig, fr, levels, canvas=prior)
        else:
            _want[key] = oracle.panoN(rig, fr, levels)
    return _want[key]


def make_pano(hm, kind, levels, rig):
    cls = {"2": hm.StitchPano, "3": hm.StitchPano3, "N": hm.StitchPanoN}[kind]
    p = cls(1, levels, rig)
    assert p.status == 0, p.status
    return p


def process(pano, kind, ins, canvas, stream=None):
    if kind == "2":
        return pano.process(ins[0], ins[1], canvas, stream)
    if kind == "3":
        return pano.process(ins[0], ins[1], ins[2], canvas, stream=stream)
    return pano.process(ins, canvas, stream)


def release(torch):
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def assert_same(got, want, what):
    diff = got != want
    assert not diff.any(), f"{what}: {int(diff.any(-1).sum())} pixels differ"


# ---------------------------------------------------------------------------
# 2 x 5312x2988 -> 8500x3300, levels 6, launch shapes
# ---------------------------------------------------------------------------
# (remap px, tail base, overlap, XCD bands, rows, canvas grid, fused down[, fused
# levels 1 and 0]): settings the autotuner picks on this workload, the fused
# up-pass levels after each tail base, the fused remap+down launch, the persistent
# canvas grids (each workgroup walks several tiles), and a shape far from the default.
CFG1_TUNINGS = [
    (4, 3, 2, 0, 2, 0, 0, 0),
    (4, 3, 2, 1, 2, 0, 0, 1),
    (4, 3, 3, 0, 1, 0, 0, 1),
    (4, 2, 3, 1, 2, 0, 1, 1),
    (4, 2, 2, 1, 2, 0, 0, 1),
    (4, 0, 2, 0, 2, 0, 0, 1),
    (4, 2, 1, 1, 1, 0, 1),
    (4, 3, 2, 1, 2, 256, 0),
    (8, 0, 0, 0, 4, 2048, 0),
    (2, 1, 1, 1, 1, 512, 1),
]


def test_cfg1_levels6_launch_shapes_bit_exact(hm, torch_gpu, oracle):
    torch = torch_gpu
    name = "2img_5312x2988_8500x3300_L6"
    rig, fr = served_rig(name), served_frames(name, 0)
    want = oracle_of(oracle, name, 0)
    pano = make_pano(hm, "2", 6, rig)
    dev = [torch.from_numpy(f).cuda() for f in fr]
    for t in CFG1_TUNINGS:
        print(f"8500x3300 L=6 tuning {t}", flush=True)
        pano.set_tuning(*t)
        canvas = torch.full(want.shape, 41, dtype=torch.uint8, device="cuda")
        assert pano.process(dev[0], dev[1], canvas) == 0
        torch.cuda.synchronize()
        assert_same(canvas.cpu().numpy(), want, f"tuning {t}")
    tuned, _ = pano.autotune(dev, canvas, frames=5)
    print(f"autotuned {tuned}", flush=True)
    canvas.fill_(41)
    assert pano.process(dev[0], dev[1], canvas) == 0
    torch.cuda.synchronize()
    assert_same(canvas.cpu().numpy(), want, f"autotuned {tuned}")
    del pano, canvas, dev
    release(torch)


# ---------------------------------------------------------------------------
# 8 x 3840x2160 -> 14000x2600, levels 6
# ---------------------------------------------------------------------------
def test_cfg2_n8_levels6_bit_exact(hm, torch_gpu, oracle):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.config_rig("cfg2_n8")
    frames = frames_for(rig, 1, 47)
    print("8 x 3840x2160 -> 14000x2600 L=6: GPU", flush=True)
    pano = make_pano(hm, "N", 6, rig)
    canvas = pano.alloc_canvas()
    canvas.fill_(77)
    assert pano.process([torch.from_numpy(f).cuda() for f in frames], canvas) == 0
    torch.cuda.synchronize()
    got = canvas.cpu().numpy()
    print("8 x 3840x2160 -> 14000x2600 L=6: oracle", flush=True)
    assert_same(got, oracle.panoN(rig, frames, 6), "14000x2600 L=6")
    del pano, canvas
    release(torch)


# ---------------------------------------------------------------------------
# the 16-frame mixed batch as one hmBatch pass
# ---------------------------------------------------------------------------
def test_cfg3_mixed_batch_full_size_every_frame_vs_oracle(hm, torch_gpu, oracle):
    import hm_synth

    torch = torch_gpu
    calls, meta = [], []
    for name, count in hm_synth.BATCH_MIX:
        kind, levels, _, n_sets = hm_synth.served_config(name)
        rig = served_rig(name)
        for j in range(count):
            k = j % n_sets
            pano = make_pano(hm, kind, levels, rig)
            assert pano.prepare() == 0
            ins = [torch.from_numpy(f).cuda() for f in served_frames(name, k)]
            canvas = pano.alloc_canvas()
            calls.append((pano, ins, canvas))
            meta.append((name, k))
    print(f"mixed batch: {len(calls)} frames", flush=True)
    assert len(calls) == 16
    batch = hm.Batch(calls)
    grid, _ = batch.autotune(passes=2)
    print(f"mixed batch: grid {grid}, {batch.info()}", flush=True)
    for replay in range(2):
        for _, _, c in calls:
            c.fill_(5 + replay)
        torch.cuda.synchronize()
        assert batch.run(use_graph=True) == 0
        torch.cuda.synchronize()
        for (name, k), (_, _, c) in zip(meta, calls):
            prior = np.full(c.shape, 5 + replay, np.uint8)
            want = oracle_of(oracle, name, k, prior if hm_synth.served_config(name)[0] == "3" else None)
            assert_same(c.cpu().numpy(), want, f"replay {replay}: {name} input set {k}")
    batch.close()
    del calls, batch
    release(torch)


# ---------------------------------------------------------------------------
# the 32-rig served stream through hmServer
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("batch", [0, 8])
def test_cfg4_served_32_rigs_each_frame_equals_alone(hm, torch_gpu, oracle, batch):
    import hm_synth

    torch = torch_gpu
    rng = np.random.default_rng(29)
    rigs = []  # (name, kind, pano, input sets, canvases)
    for name, kind, _geo, levels, count, n_sets in hm_synth.SERVED_CONFIGS:
        print(f"served: {count} rigs of {name}", flush=True)
        rig = served_rig(name)
        sets = [[torch.from_numpy(f).cuda() for f in served_frames(name, k)] for k in range(n_sets)]
        tuned = None
        for _ in range(count):
            pano = make_pano(hm, kind, levels, rig)
            assert pano.prepare() == 0
            canvases = [pano.alloc_canvas(), pano.alloc_canvas()]
            if tuned is None:
                tuned, _ = pano.autotune(sets[0], canvases[0], frames=3)
            pano.set_tuning(**tuned)
            rigs.append((name, kind, pano, sets, canvases))
    assert len(rigs) == 32
    # alone: each (config, input set) once through process, on a fresh object of its config
    alone = {}
    for name, kind, pano, sets, canvases in rigs:
        for k in range(len(sets)):
            if (name, k) in alone:
                continue
            c = canvases[0]
            c.fill_(3)
            torch.cuda.synchronize()
            assert process(pano, kind, sets[k], c) == 0
            torch.cuda.synchronize()
            alone[name, k] = c.cpu().numpy()
    # one frame per config against the oracle
    for name, kind, _geo, _levels, _count, _sets in hm_synth.SERVED_CONFIGS:
        prior = np.full(alone[name, 0].shape, 3, np.uint8) if kind == "3" else None
        assert_same(alone[name, 0], oracle_of(oracle, name, 0, prior), f"{name} alone vs oracle")
    # the mix: two frames per rig (each into its own canvas), in a shuffled order, over 4 streams
    jobs = [(ri, r) for r in range(2) for ri in range(len(rigs))]
    rng.shuffle(jobs)
    for _, _, _, _, canvases in rigs:
        for c in canvases:
            c.fill_(3)
    torch.cuda.synchronize()
    server = hm.Server(n_streams=4, device_budget_bytes=200 << 30, use_graphs=True, batch=batch)
    ids = [server.add_rig(r[2]) for r in rigs]
    for tag, (ri, r) in enumerate(jobs):
        name, kind, pano, sets, canvases = rigs[ri]
        assert server.submit(ids[ri], sets[r % len(sets)], canvases[r], tag=tag) == 0
    assert server.wait_idle() == 0
    res = server.poll(4096)
    assert sorted(x["tag"] for x in res) == list(range(len(jobs)))
    assert all(x["status"] == 0 for x in res)
    for ri, (name, kind, pano, sets, canvases) in enumerate(rigs):
        for r in range(2):
            assert_same(canvases[r].cpu().numpy(), alone[name, r % len(sets)], f"rig {ri} ({name}) frame {r}")
    st = server.stats()
    print(f"served stats: {st}", flush=True)
    if batch:
        assert st["batched_passes"] > 0
    server.close()
    del rigs, server
    release(torch)
