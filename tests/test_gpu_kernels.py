"""GPU parity of the kernel-tier entry points (remap, ROI copy, blend contexts)
against the CPU oracle and the reference's blend known answers."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")
KAT = json.load(open(GOLDEN))
KIND = {"2": 2, "3": 3, "N": 0}


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rand_images(rng, n, b, h, w, c, holes=True):
    imgs = []
    for i in range(n):
        im = rng.integers(0, 256, (b, h, w, c)).astype(np.float32)
        if c == 4:
            im[..., 3] = 255
            if holes:
                # transparent regions that differ per image, as remapped footprints do
                y0, x0 = rng.integers(0, max(1, h // 2)), rng.integers(0, max(1, w // 2))
                im[:, y0: y0 + h // 3, x0: x0 + w // 3] = 0
                im[:, :, : i * 3] = 0
        imgs.append(im)
    return imgs


# ---------------------------------------------------------------------------
# blend contexts
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
def test_blend_known_answers_gpu(hm, torch_gpu, case):
    torch = torch_gpu
    c, n = case["c"], len(case["images"])
    imgs = [dev(torch, np.float32(v).reshape(1, 1, 1, c)) for v in case["images"]]
    mask = dev(torch, np.float32(case["mask"]))
    out = torch.zeros((1, 1, 1, c), dtype=torch.float32, device="cuda")
    ctx = hm.BlendContext(KIND[case["kind"]], n, 1, 1, case["levels"], 1, c)
    ctx.run(imgs, mask, out)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy().reshape(-1), case["expected"], atol=case["tol"], rtol=0)


@pytest.mark.parametrize("kind,n", [(2, 2), (3, 3), (0, 2), (0, 3), (0, 5), (0, 8)])
@pytest.mark.parametrize("c", [4, 3])
@pytest.mark.parametrize("shape,levels", [((37, 53), 5), ((64, 48), 3), ((5, 3), 6), ((1, 1), 1), ((131, 97), 8),
                                          ((70, 200), 2)])
@pytest.mark.parametrize("debug", [False, True], ids=["fused", "per_level"])
def test_blend_context_matches_oracle(hm, torch_gpu, oracle, kind, n, c, shape, levels, debug):
    torch = torch_gpu
    rng = np.random.default_rng(hash((kind, n, c, shape, levels)) & 0xFFFF)
    h, w = shape
    b = 2
    imgs = rand_images(rng, n, b, h, w, c)
    if kind == 2:
        mask = rng.uniform(0, 1, (h, w)).astype(np.float32)
        mask[: h // 2] = np.round(mask[: h // 2])
    else:
        mask = rng.uniform(0, 1, (h, w, n)).astype(np.float32)
        mask[: h // 2] = (mask[: h // 2] == mask[: h // 2].max(-1, keepdims=True)).astype(np.float32)
        mask[0, 0] = 0
    want, d = oracle.blend(kind, imgs, mask, levels, dump=True)
    ctx = hm.BlendContext(kind, n, w, h, levels, b, c, debug=debug)
    assert ctx.levels() == d.sizes
    out = torch.zeros((b, h, w, c), dtype=torch.float32, device="cuda")
    dimgs = [dev(torch, i) for i in imgs]
    ctx.run(dimgs, dev(torch, mask), out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_blend_output_may_alias_first_image(hm, torch_gpu, oracle):
    """The 2-image pano passes full1 as both input and output (cudaPano.inl:384-392)."""
    torch = torch_gpu
    rng = np.random.default_rng(4)
    imgs = rand_images(rng, 2, 1, 33, 41, 4)
    mask = rng.uniform(0, 1, (33, 41)).astype(np.float32)
    want = oracle.blend(2, imgs, mask, 4)
    d = [dev(torch, i) for i in imgs]
    ctx = hm.BlendContext(2, 2, 41, 33, 4, 1, 4)
    ctx.run(d, dev(torch, mask), d[0])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d[0].cpu().numpy(), want)


# ---------------------------------------------------------------------------
# remap / copy
# ---------------------------------------------------------------------------
def rand_maps(rng, mw, mh, sw, sh):
    mx = rng.integers(0, sw + 3, (mh, mw)).astype(np.uint16)
    my = rng.integers(0, sh + 3, (mh, mw)).astype(np.uint16)
    sent = rng.uniform(size=(mh, mw)) < 0.1
    mx[sent] = 65535
    my[sent] = 65535
    mx[rng.uniform(size=(mh, mw)) < 0.05] = 65535  # x sentinel with a valid y
    return mx, my


@pytest.mark.parametrize("c,dst_float", [(4, False), (4, True), (3, False), (3, True)])
@pytest.mark.parametrize("no_unmapped", [False, True])
def test_remap_ex_offset_matches_oracle(hm, torch_gpu, oracle, c, dst_float, no_unmapped):
    torch = torch_gpu
    rng = np.random.default_rng(c * 10 + dst_float + 2 * no_unmapped)
    src = rng.integers(0, 256, (2, 20, 30, c), dtype=np.uint8)
    mx, my = rand_maps(rng, 25, 17, 30, 20)
    dtype = np.float32 if dst_float else np.uint8
    dst0 = np.full((2, 22, 40, c), 9, dtype)
    want = dst0.copy()
    oracle.remap_offset(src, want, mx, my, 7, -3, no_unmapped_write=no_unmapped)
    d = dev(torch, dst0)
    st = hm.HM_PIX_U8C4 if c == 4 else hm.HM_PIX_U8C3
    dt = (hm.HM_PIX_F32C4 if c == 4 else hm.HM_PIX_F32C3) if dst_float else st
    rc = hm.remap_ex_offset(dev(torch, src), st, d, dt, dev(torch, mx), dev(torch, my), 2, 25, 17, 7, -3, no_unmapped)
    assert rc == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)
    # ROI form, plus its validation
    want2 = dst0.copy()
    oracle.remap_offset(src, want2, mx, my, 2, 1, no_unmapped_write=no_unmapped, roi=(3, 2, 11, 9))
    d2 = dev(torch, dst0)
    rc = hm.remap_ex_offset(dev(torch, src), st, d2, dt, dev(torch, mx), dev(torch, my), 2, 25, 17, 2, 1,
                            no_unmapped, roi=(3, 2, 11, 9))
    assert rc == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d2.cpu().numpy(), want2)
    assert hm.remap_ex_offset(dev(torch, src), st, d2, dt, dev(torch, mx), dev(torch, my), 2, 25, 17, 2, 1,
                              no_unmapped, roi=(20, 2, 11, 9)) != 0


@pytest.mark.parametrize("c", [4, 3])
def test_remap_dest_map_and_hard_seam_n(hm, torch_gpu, oracle, c):
    torch = torch_gpu
    rng = np.random.default_rng(c)
    t = hm.HM_PIX_U8C4 if c == 4 else hm.HM_PIX_U8C3
    src = rng.integers(0, 256, (2, 20, 30, c), dtype=np.uint8)
    mx, my = rand_maps(rng, 25, 17, 30, 20)
    dm = rng.integers(0, 3, (22, 40)).astype(np.uint8)
    dst0 = np.full((2, 22, 40, c), 5, np.uint8)
    want = dst0.copy()
    oracle.remap_dest_map(src, want, mx, my, 4, 2, 1, dm)
    d = dev(torch, dst0)
    assert hm.remap_dest_map(dev(torch, src), d, t, dev(torch, mx), dev(torch, my), 1, dev(torch, dm), 2, 25, 17, 4,
                             2) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)
    # fused N-image hard seam
    srcs = [rng.integers(0, 256, (2, 20, 30, c), dtype=np.uint8) for _ in range(3)]
    maps = [rand_maps(rng, 25, 17, 30, 20) for _ in range(3)]
    offs = [(0, 0), (8, 3), (15, 5)]
    want = dst0.copy()
    oracle.remap_hard_seam_n(srcs, [m[0] for m in maps], [m[1] for m in maps], offs, dm, want)
    d = dev(torch, dst0)
    assert hm.remap_hard_seam_n([dev(torch, s) for s in srcs], [dev(torch, m[0]) for m in maps],
                                [dev(torch, m[1]) for m in maps], offs, [(25, 17)] * 3, dev(torch, dm), d, t, 2) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


@pytest.mark.parametrize("sc,sf,dc,df", [(4, True, 4, False), (4, False, 4, True), (3, True, 3, False),
                                         (3, True, 4, True), (4, True, 3, False), (3, False, 4, True)])
def test_copy_roi_and_make_full(hm, torch_gpu, oracle, sc, sf, dc, df):
    torch = torch_gpu
    rng = np.random.default_rng(sc + 2 * dc + 4 * sf + 8 * df)
    if sf:
        src = rng.uniform(-20, 280, (2, 12, 15, sc)).astype(np.float32)
    else:
        src = rng.integers(0, 256, (2, 12, 15, sc), dtype=np.uint8)
    dt = np.float32 if df else np.uint8
    dst0 = np.full((2, 10, 13, dc), 3, dt)

    def code(c, f):
        return {(3, False): hm.HM_PIX_U8C3, (4, False): hm.HM_PIX_U8C4, (3, True): hm.HM_PIX_F32C3,
                (4, True): hm.HM_PIX_F32C4}[(c, f)]

    want = dst0.copy()
    oracle.copy_roi(src, 9, 8, 4, 2, want, 5, 3)
    d = dev(torch, dst0)
    assert hm.copy_roi(dev(torch, src), code(sc, sf), 9, 8, 4, 2, d, code(dc, df), 5, 3, 2) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)
    want = np.zeros_like(dst0)
    oracle.copy_roi(src, 9, 8, 4, 2, want, 5, 3)
    d = dev(torch, dst0)
    assert hm.make_full(dev(torch, src), code(sc, sf), 9, 8, 4, 2, 5, 3, 2, d, code(dc, df)) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


# ---------------------------------------------------------------------------
# float, half and 1-channel pixel types (src/cuda/cudaRemap.cu:507-554,
# cudaMakeFull.cu:291-313)
# ---------------------------------------------------------------------------
def _ptype(hm, c, kind):
    return {("u8", 1): hm.HM_PIX_U8C1, ("u8", 3): hm.HM_PIX_U8C3, ("u8", 4): hm.HM_PIX_U8C4,
            ("f32", 1): hm.HM_PIX_F32C1, ("f32", 3): hm.HM_PIX_F32C3, ("f32", 4): hm.HM_PIX_F32C4,
            ("f16", 1): hm.HM_PIX_F16C1, ("f16", 3): hm.HM_PIX_F16C3, ("f16", 4): hm.HM_PIX_F16C4}[(kind, c)]


def _rand_px(rng, shape, kind):
    if kind == "u8":
        return rng.integers(0, 256, shape, dtype=np.uint8)
    v = (rng.uniform(-20, 290, shape)).astype(np.float32)
    return v.astype(np.float16) if kind == "f16" else v


@pytest.mark.parametrize("c,kind", [(1, "f32"), (3, "f32"), (4, "f32"), (1, "u8"), (4, "f16")])
def test_remap_ex_offset_float_and_one_channel_sources(hm, torch_gpu, oracle, c, kind):
    """float1/3/4 -> same (and uchar1, half4): exact copies; defaults with alpha 0 / 255."""
    torch = torch_gpu
    rng = np.random.default_rng(c * 7 + len(kind))
    src = _rand_px(rng, (2, 20, 30, c), kind)
    mx, my = rand_maps(rng, 25, 17, 30, 20)
    dst0 = np.full((2, 22, 40, c), 9, src.dtype)
    want = dst0.copy()
    oracle.remap_offset(src, want, mx, my, 7, -3)
    d = dev(torch, dst0)
    t = _ptype(hm, c, kind)
    assert hm.remap_ex_offset(dev(torch, src), t, d, t, dev(torch, mx), dev(torch, my), 2, 25, 17, 7, -3, False) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


def test_remap_ex_offset_refuses_pairs_not_instantiated(hm, torch_gpu):
    torch = torch_gpu
    src = torch.zeros((1, 4, 4, 4), dtype=torch.float32, device="cuda")
    dst = torch.zeros((1, 4, 4, 4), dtype=torch.uint8, device="cuda")
    m = torch.zeros((4, 4), dtype=torch.int16, device="cuda")
    assert hm.remap_ex_offset(src, hm.HM_PIX_F32C4, dst, hm.HM_PIX_U8C4, m, m, 1, 4, 4, 0, 0, False) != 0
    assert hm.remap_ex_offset(src, hm.HM_PIX_F32C4, src, hm.HM_PIX_F32C3, m, m, 1, 4, 4, 0, 0, False) != 0


@pytest.mark.parametrize("c,kind", [(1, "u8"), (1, "f32"), (1, "f16"), (3, "f32"), (4, "f32")])
def test_remap_dest_map_other_types(hm, torch_gpu, oracle, c, kind):
    """uchar1, float1, __half, float3, float4 dest-map remaps."""
    torch = torch_gpu
    rng = np.random.default_rng(11 * c + len(kind))
    src = _rand_px(rng, (2, 20, 30, c), kind)
    mx, my = rand_maps(rng, 25, 17, 30, 20)
    dm = rng.integers(0, 3, (22, 40)).astype(np.uint8)
    dst0 = np.full((2, 22, 40, c), 5, src.dtype)
    want = dst0.copy()
    oracle.remap_dest_map(src, want, mx, my, 4, 2, 1, dm)
    d = dev(torch, dst0)
    assert hm.remap_dest_map(dev(torch, src), d, _ptype(hm, c, kind), dev(torch, mx), dev(torch, my), 1,
                             dev(torch, dm), 2, 25, 17, 4, 2) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


@pytest.mark.parametrize("c", [3, 4])
def test_remap_hard_seam_n_float(hm, torch_gpu, oracle, c):
    torch = torch_gpu
    rng = np.random.default_rng(40 + c)
    dm = rng.integers(0, 4, (22, 40)).astype(np.uint8)  # label 3: no image
    srcs = [_rand_px(rng, (2, 20, 30, c), "f32") for _ in range(3)]
    maps = [rand_maps(rng, 25, 17, 30, 20) for _ in range(3)]
    offs = [(0, 0), (8, 3), (15, 5)]
    dst0 = np.full((2, 22, 40, c), 5, np.float32)
    want = dst0.copy()
    oracle.remap_hard_seam_n(srcs, [m[0] for m in maps], [m[1] for m in maps], offs, dm, want)
    d = dev(torch, dst0)
    assert hm.remap_hard_seam_n([dev(torch, s) for s in srcs], [dev(torch, m[0]) for m in maps],
                                [dev(torch, m[1]) for m in maps], offs, [(25, 17)] * 3, dev(torch, dm), d,
                                _ptype(hm, c, "f32"), 2) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), want)


@pytest.mark.parametrize("s,d", [((3, "f16"), (3, "u8")), ((3, "f16"), (4, "u8")), ((4, "f16"), (4, "u8")),
                                 ((4, "f16"), (3, "u8")), ((3, "f32"), (4, "f16")), ((4, "f16"), (3, "f32")),
                                 ((3, "u8"), (4, "f16")), ((4, "u8"), (3, "f16")), ((3, "f32"), (3, "f16"))])
def test_copy_roi_and_make_full_half_types(hm, torch_gpu, oracle, s, d):
    """half3/half4 copies and make_full variants: to uint8 clamps and truncates, to half rounds
    to nearest even, 3 -> 4 sets alpha 255, 4 -> 3 drops it."""
    torch = torch_gpu
    rng = np.random.default_rng(hash((s, d)) & 0xFFFF)
    src = _rand_px(rng, (2, 15, 21, s[0]), s[1])
    ddt = {"u8": np.uint8, "f32": np.float32, "f16": np.float16}[d[1]]
    dst0 = np.full((2, 18, 26, d[0]), 3, ddt)
    want = dst0.copy()
    oracle.copy_roi(src, 13, 9, 4, 3, want, 6, 5)
    dd = dev(torch, dst0)
    assert hm.copy_roi(dev(torch, src), _ptype(hm, *s), 13, 9, 4, 3, dd, _ptype(hm, *d), 6, 5, 2) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint8), want.view(np.uint8))
    # simple_make_full_batch: the destination cleared, then the copy
    want = np.zeros_like(dst0)
    oracle.copy_roi(src, 13, 9, 4, 3, want, 6, 5)
    dd = dev(torch, dst0)
    assert hm.make_full(dev(torch, src), _ptype(hm, *s), 13, 9, 4, 3, 6, 5, 2, dd, _ptype(hm, *d)) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint8), want.view(np.uint8))
