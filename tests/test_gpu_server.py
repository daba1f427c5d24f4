"""The served form (hmServer*): many rigs of every pano class in flight on one
GPU, hipGraph replay, admission by device memory.  Every served frame must be
bit-identical to the same frame stitched alone (the oracle)."""
import numpy as np
import pytest

from conftest import frames_for

pytestmark = pytest.mark.gpu

HIP_ERROR_OUT_OF_MEMORY = 2


def to_dev(torch, frames):
    return [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]


def make_rigs(hm):
    """(name, pano, rig, levels, kind, fused) for a mix of classes and levels."""
    import hm_synth

    out = []
    r2 = hm_synth.small_rig(seed=11)
    out.append(("pano2_L4", hm.StitchPano(1, 4, r2), r2, 4, "2", True))
    r2b = hm_synth.small_rig(seed=12)
    out.append(("pano2_L0", hm.StitchPano(1, 0, r2b), r2b, 0, "2", True))
    rn = hm_synth.small_rig(n=3, seed=13, src=(120, 56), canvas=(330, 64))
    out.append(("panoN3_L3", hm.StitchPanoN(1, 3, rn), rn, 3, "N", True))
    rn5 = hm_synth.small_rig(n=5, seed=14, src=(120, 56), canvas=(510, 64))
    out.append(("panoN5_L0", hm.StitchPanoN(1, 0, rn5), rn5, 0, "N", True))
    r3 = hm_synth.small_rig(n=3, seed=15, src=(120, 56), canvas=(330, 64))
    out.append(("pano3_L2_fused", hm.StitchPano3(1, 2, r3), r3, 2, "3", True))
    r3b = hm_synth.small_rig(n=3, seed=16, src=(120, 56), canvas=(330, 64))
    out.append(("pano3_L3_unfused", hm.StitchPano3(1, 3, r3b), r3b, 3, "3", False))
    return out


def oracle_out(oracle, kind, rig, frames, levels, fused, prior):
    if kind == "2":
        return oracle.pano2(rig, frames[0], frames[1], levels)
    if kind == "N":
        return oracle.panoN(rig, frames, levels)
    return oracle.pano3(rig, frames, levels, fused=fused, canvas=prior)


@pytest.mark.parametrize("use_graphs", [True, False])
def test_server_mixed_rigs_match_alone(hm, torch_gpu, oracle, use_graphs):
    torch = torch_gpu
    rigs = make_rigs(hm)
    server = hm.Server(n_streams=4, use_graphs=use_graphs)
    ids = [server.add_rig(p) for (_, p, *_rest) in rigs]
    inputs = {}  # (rig index, input set) -> (host frames, device tensors)
    for ri, (_, _, rig, *_r) in enumerate(rigs):
        for s in range(2):
            fr = frames_for(rig, 1, 100 * ri + s)
            inputs[ri, s] = (fr, to_dev(torch, fr))
    rounds = 4
    jobs = []
    for rnd in range(rounds):
        for ri, (name, pano, rig, levels, kind, fused) in enumerate(rigs):
            w, h = pano.canvas_size()
            prior = np.full((1, h, w, 4), 9 + rnd, np.uint8)
            jobs.append((ri, rnd % 2, torch.from_numpy(prior).cuda(), prior))
    # the server's streams do not wait for torch's: every buffer is ready first
    torch.cuda.synchronize()
    for tag, (ri, s, canvas, prior) in enumerate(jobs):
        assert server.submit(ids[ri], inputs[ri, s][1], canvas, tag=tag, fused=rigs[ri][5]) == 0
    assert server.wait_idle() == 0
    res = server.poll()
    assert sorted(r["tag"] for r in res) == list(range(len(jobs)))
    for r in res:
        assert r["status"] == 0
        assert r["t_submit_ns"] <= r["t_start_ns"] <= r["t_done_ns"]
    for tag, (ri, s, canvas, prior) in enumerate(jobs):
        name, pano, rig, levels, kind, fused = rigs[ri]
        want = oracle_out(oracle, kind, rig, inputs[ri, s][0], levels, fused, prior)
        np.testing.assert_array_equal(canvas.cpu().numpy(), want, err_msg=f"{name} round {tag}")
    st = server.stats()
    if use_graphs:
        assert st["graphs_captured"] == len(jobs) and st["graph_launches"] == len(jobs)
    else:
        assert st["eager_launches"] == len(jobs)
    server.close()


@pytest.mark.parametrize("use_graphs", [True, False])
@pytest.mark.parametrize("max_batch", [2, 8])
def test_server_batched_passes_match_alone(hm, torch_gpu, oracle, use_graphs, max_batch):
    """Continuous batching: requests of every class queued together run as batched passes
    of up to max_batch rigs (replayed as graphs when a composition comes back); a float
    pipeline rig, which the batched kernels do not carry, runs alone in the same stream of
    requests.  Every frame equals the oracle's, rounds after rounds."""
    import hm_synth

    torch = torch_gpu
    rigs = make_rigs(hm)
    rf = hm_synth.small_rig(seed=17)
    fpano = hm.StitchPano(1, 3, rf, float_pipeline=True)
    server = hm.Server(n_streams=2, use_graphs=use_graphs, batch=max_batch)
    ids = [server.add_rig(p) for (_, p, *_rest) in rigs]
    fid = server.add_rig(fpano)
    inputs = {}
    for ri, (_, _, rig, *_r) in enumerate(rigs):
        for s in range(2):
            fr = frames_for(rig, 1, 300 + 100 * ri + s)
            inputs[ri, s] = (fr, to_dev(torch, fr))
    ffr = [f.astype(np.float32) + 0.25 for f in frames_for(rf, 1, 77)]
    fdev = to_dev(torch, ffr)
    fcanvas = torch.zeros((1, fpano.canvas_size()[1], fpano.canvas_size()[0], 4), dtype=torch.float32,
                          device="cuda")
    canvases = {}
    for ri, (_, pano, *_r) in enumerate(rigs):
        w, h = pano.canvas_size()
        for s in range(2):
            canvases[ri, s] = torch.zeros((1, h, w, 4), dtype=torch.uint8, device="cuda")
    rounds = 6
    tag = 0
    for rnd in range(rounds):
        s = rnd % 2
        priors = {}
        for ri in range(len(rigs)):
            priors[ri] = np.full(canvases[ri, s].shape, 9 + rnd, np.uint8)
            canvases[ri, s].copy_(torch.from_numpy(priors[ri]).cuda())
        torch.cuda.synchronize()
        jobs = {}
        reqs = []  # one tick of every rig, queued at once so that passes form
        for ri, (name, pano, rig, levels, kind, fused) in enumerate(rigs):
            reqs.append((ids[ri], inputs[ri, s][1], canvases[ri, s], tag, fused))
            jobs[tag] = ri
            tag += 1
        reqs.insert(3, (fid, fdev, fcanvas, tag))
        jobs[tag] = -1
        tag += 1
        assert server.submit_many(reqs) == 0
        assert server.wait_idle() == 0
        res = server.poll()
        assert sorted(r["tag"] for r in res) == sorted(jobs)
        assert all(r["status"] == 0 for r in res)
        for ri, (name, pano, rig, levels, kind, fused) in enumerate(rigs):
            want = oracle_out(oracle, kind, rig, inputs[ri, s][0], levels, fused, priors[ri])
            np.testing.assert_array_equal(canvases[ri, s].cpu().numpy(), want, err_msg=f"{name} round {rnd}")
        np.testing.assert_array_equal(fcanvas.cpu().numpy(), oracle.pano2(rf, ffr[0], ffr[1], 3),
                                      err_msg=f"float rig round {rnd}")
    st = server.stats()
    print(st)
    # every tick: the six carried rigs in passes of up to max_batch, the float rig alone; the
    # compositions of rounds 0/2/4 and 1/3/5 repeat, so with graphs rounds 4 and 5 replay them
    per_tick = -(-len(rigs) // max_batch)
    assert st["batched_passes"] == rounds * per_tick and st["batched_frames"] == rounds * len(rigs)
    if use_graphs:
        assert st["pass_graphs_captured"] == 2 * per_tick and st["pass_graph_launches"] == 4 * per_tick
    else:
        assert st["pass_graph_launches"] == 0
    bad = (ids[0], inputs[0, 0][1], canvases[0, 0], 999)
    assert server.submit_many([bad, (ids[0] + 100, inputs[0, 0][1], canvases[0, 0], 1000)]) != 0
    assert server.stats()["queued"] == 0  # all or nothing
    server.close()


def test_server_batching_same_structure(hm, torch_gpu, oracle):
    """Policy 'same structure': only rigs whose pipelines launch the same kernel sequence share
    a pass (here two 2-image L4 rigs of different canvas sizes, and two 3-image N-class L3
    rigs); a rig of another structure (2-image L0) runs alone."""
    import hm_synth

    torch = torch_gpu
    specs = [
        ("2", 4, hm_synth.small_rig(seed=51)),
        ("N", 3, hm_synth.small_rig(n=3, seed=52, src=(120, 56), canvas=(330, 64))),
        ("2", 0, hm_synth.small_rig(seed=53)),
        ("2", 4, hm_synth.small_rig(seed=54, src=(170, 72), canvas=(440, 80))),
        ("N", 3, hm_synth.small_rig(n=3, seed=55, src=(100, 48), canvas=(280, 56))),
    ]
    rigs = []
    for kind, levels, rig in specs:
        pano = hm.StitchPano(1, levels, rig) if kind == "2" else hm.StitchPanoN(1, levels, rig)
        assert pano.status == 0
        rigs.append((kind, levels, rig, pano))
    assert rigs[0][3].canvas_size() != rigs[3][3].canvas_size()
    server = hm.Server(n_streams=2, use_graphs=True, batch=8, same_structure=True)
    ids = [server.add_rig(p) for (*_r, p) in rigs]
    for rnd in range(3):
        frames = [frames_for(rig, 1, 600 + 10 * ri + rnd) for ri, (_, _, rig, _) in enumerate(rigs)]
        devs = [to_dev(torch, f) for f in frames]
        canvases = [p.alloc_canvas() for (*_r, p) in rigs]
        torch.cuda.synchronize()
        assert server.submit_many([(ids[i], devs[i], canvases[i], i) for i in range(len(rigs))]) == 0
        assert server.wait_idle() == 0
        assert all(r["status"] == 0 for r in server.poll())
        for i, (kind, levels, rig, _) in enumerate(rigs):
            want = oracle.pano2(rig, frames[i][0], frames[i][1], levels) if kind == "2" else \
                oracle.panoN(rig, frames[i], levels)
            np.testing.assert_array_equal(canvases[i].cpu().numpy(), want, err_msg=f"rig {i} round {rnd}")
    st = server.stats()
    assert st["batched_passes"] == 2 * 3 and st["batched_frames"] == 4 * 3
    server.close()


def test_server_remove_rig_while_batched_passes_run(hm, torch_gpu, oracle):
    """A rig leaves while passes of the other rigs (captured as graphs) are queued and running:
    only the passes that ran the leaving rig's recordings are dropped, the others keep replaying,
    and every frame of the remaining rigs matches the oracle."""
    import hm_synth

    torch = torch_gpu
    rigs = [hm_synth.small_rig(seed=70 + k) for k in range(4)]
    panos = [hm.StitchPano(1, 3, r) for r in rigs]
    server = hm.Server(n_streams=2, use_graphs=True, batch=4, same_structure=True)
    ids = [server.add_rig(p) for p in panos]
    frames = [frames_for(r, 1, 800 + k) for k, r in enumerate(rigs)]
    devs = [to_dev(torch, f) for f in frames]
    canvases = [p.alloc_canvas() for p in panos]
    torch.cuda.synchronize()
    tag = 0
    for rnd in range(3):  # compositions of all four rigs, seen often enough to be captured
        assert server.submit_many([(ids[k], devs[k], canvases[k], tag + k) for k in range(4)]) == 0
        tag += 4
        assert server.wait_idle() == 0
    # rigs 1..3 keep submitting while rig 0 leaves
    for rnd in range(4):
        assert server.submit_many([(ids[k], devs[k], canvases[k], tag + k) for k in range(1, 4)]) == 0
        tag += 3
        if rnd == 1:
            server.remove_rig(ids[0])
    assert server.wait_idle() == 0
    res = server.poll(4096)
    assert all(r["status"] == 0 for r in res)
    for k in range(1, 4):
        np.testing.assert_array_equal(canvases[k].cpu().numpy(), oracle.pano2(rigs[k], frames[k][0], frames[k][1], 3),
                                      err_msg=f"rig {k}")
    st = server.stats()
    assert st["n_rigs"] == 3 and st["pass_graph_launches"] > 0
    server.close()


def test_server_graph_replays_with_eager_work_between(hm, torch_gpu, oracle):
    """One buffer set, replayed several times with new input content written in between."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=21)
    pano = hm.StitchPano(1, 5, rig)
    server = hm.Server(n_streams=2, use_graphs=True)
    rid = server.add_rig(pano)
    dev = to_dev(torch, frames_for(rig, 1, 0))
    canvas = pano.alloc_canvas()
    torch.cuda.synchronize()
    for k in range(5):
        fr = frames_for(rig, 1, 50 + k)
        for d, f in zip(dev, fr):
            d.copy_(torch.from_numpy(f).cuda())  # eager work on torch's stream
        torch.cuda.synchronize()
        assert server.submit(rid, dev, canvas, tag=k) == 0
        assert server.wait_idle() == 0
        scratch = torch.ones(1 << 20, device="cuda").sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert float(scratch) == float(1 << 20)
        np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.pano2(rig, fr[0], fr[1], 5), err_msg=f"replay {k}")
    st = server.stats()
    assert st["graphs_captured"] == 1 and st["graph_launches"] == 5
    server.close()


def test_server_admission_by_device_budget(hm, torch_gpu):
    import hm_synth

    rig = hm_synth.small_rig(seed=31)
    a = hm.StitchPano(1, 4, rig)
    b = hm.StitchPano(1, 4, rig)
    assert a.prepare() == 0
    need = a.device_bytes()
    assert need > 0
    server = hm.Server(n_streams=2, device_budget_bytes=need + need // 2)
    ra = server.add_rig(a)
    with pytest.raises(hm.HmError) as ei:
        server.add_rig(b)
    assert ei.value.code == HIP_ERROR_OUT_OF_MEMORY
    assert server.stats()["n_rigs"] == 1
    server.remove_rig(ra)
    rb = server.add_rig(b)
    assert server.stats()["device_used_bytes"] == b.device_bytes()
    server.remove_rig(rb)
    server.close()


def test_server_request_errors(hm, torch_gpu, oracle):
    """Bad submissions are refused; a request whose process fails reports its status and the
    server goes on with the next one (the N-image class's errors are not sticky)."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(n=3, seed=41, src=(120, 56), canvas=(330, 64))
    pano = hm.StitchPanoN(1, 2, rig)
    server = hm.Server(n_streams=2)
    rid = server.add_rig(pano)
    fr = frames_for(rig, 1, 3)
    dev = to_dev(torch, fr)
    good = pano.alloc_canvas()
    bad = torch.zeros((1, 8, 8, 4), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    assert server.submit(rid, dev[:2], good) != 0  # wrong input count
    assert server.submit(rid + 100, dev, good) != 0  # unknown rig
    assert server.submit(rid, dev, bad, tag=1) == 0
    assert server.submit(rid, dev, good, tag=2) == 0
    server.wait_idle()
    res = {r["tag"]: r for r in server.poll()}
    assert res[1]["status"] != 0 and res[2]["status"] == 0
    np.testing.assert_array_equal(good.cpu().numpy(), oracle.panoN(rig, fr, 2))
    server.close()


def test_served_pano_refuses_tuning_and_batching(hm, torch_gpu):
    """While a server holds a pano its scheduler thread launches it and the captured graphs
    carry its launch shapes: tuning, autotuning, a batch over it and a second server are
    refused until the rig is removed."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=44)
    pano = hm.StitchPano(1, 4, rig)
    dev = to_dev(torch, frames_for(rig, 1, 2))
    canvas = pano.alloc_canvas()
    server = hm.Server(n_streams=2)
    rid = server.add_rig(pano)
    other = hm.Server(n_streams=1)
    with pytest.raises(hm.HmError):
        other.add_rig(pano)
    with pytest.raises(hm.HmError):
        pano.set_tuning(4, 3, 2)
    with pytest.raises(hm.HmError):
        pano.autotune(dev, canvas, frames=1)
    with pytest.raises(hm.HmError):
        hm.Batch([(pano, dev, canvas)])
    server.remove_rig(rid)
    pano.set_tuning(4, 3, 2)
    assert other.add_rig(pano) >= 0
    other.close()
    pano.set_tuning(4, 2, 1)
    server.close()
