"""hmBlendRunNoAlloc (the allocation-free blend: pyramid levels in the library's
static device scratch) against hmBlendContextCreate + hmBlendRun, through the
C ABI.  Two calls in a row per blend kind, the second smaller than the first,
must each equal the context path bit for bit (and the oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP_ERROR_OUT_OF_MEMORY = 2


def _inputs(rng, kind, n, b, h, w, c):
    imgs = []
    for i in range(n):
        im = rng.integers(0, 256, (b, h, w, c)).astype(np.float32)
        if c == 4:
            im[..., 3] = 255
            y0, x0 = rng.integers(0, max(1, h // 2)), rng.integers(0, max(1, w // 2))
            im[:, y0: y0 + h // 3, x0: x0 + w // 3] = 0
            im[:, :, : i * 3] = 0
        imgs.append(im)
    if kind == 2:
        mask = rng.uniform(0, 1, (h, w)).astype(np.float32)
        mask[: h // 2] = np.round(mask[: h // 2])
    else:
        mask = rng.uniform(0, 1, (h, w, n)).astype(np.float32)
        mask[: h // 2] = (mask[: h // 2] == mask[: h // 2].max(-1, keepdims=True)).astype(np.float32)
    return imgs, mask


@pytest.mark.parametrize("kind,n,c", [(2, 2, 4), (3, 3, 4), (0, 4, 4), (0, 3, 3)])
def test_no_alloc_blend_equals_context_blend(hm, torch_gpu, oracle, kind, n, c):
    torch = torch_gpu
    rng = np.random.default_rng(kind * 10 + n + c)
    # first call large (1200 x 700, 6 levels), second smaller (333 x 201, 4 levels): the
    # second must not see what the first left in the shared scratch
    for (h, w, levels, b) in [(700, 1200, 6, 1), (201, 333, 4, 2)]:
        print(f"no-alloc blend kind={kind} n={n} c={c} {w}x{h} L={levels} B={b}", flush=True)
        imgs, mask = _inputs(rng, kind, n, b, h, w, c)
        d_imgs = [torch.from_numpy(i).cuda() for i in imgs]
        d_mask = torch.from_numpy(mask).cuda()
        want = torch.full((b, h, w, c), -7.0, dtype=torch.float32, device="cuda")
        ctx = hm.BlendContext(kind, n, w, h, levels, b, c)
        ctx.run(d_imgs, d_mask, want)
        got = torch.full((b, h, w, c), -9.0, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        rc = hm.blend_run_no_alloc(kind, n, w, h, levels, b, c, d_imgs, d_mask, got)
        assert rc == 0, rc
        torch.cuda.synchronize()
        g = got.cpu().numpy()
        np.testing.assert_array_equal(g, want.cpu().numpy())
        if h * w <= 201 * 333:
            np.testing.assert_array_equal(g, oracle.blend(kind, imgs, mask, levels))


def test_no_alloc_blend_rejects_a_pyramid_larger_than_the_scratch(hm, torch_gpu):
    """8000 x 8000 at 2 levels needs 3 x 4000 x 4000 x 16 B of level 1 alone, beyond the
    128 MiB scratch: refused before any launch, and a later call still works."""
    torch = torch_gpu
    tiny = [torch.zeros((1, 1, 1, 4), dtype=torch.float32, device="cuda") for _ in range(2)]
    m = torch.zeros((1, 1), dtype=torch.float32, device="cuda")
    out = torch.zeros((1, 1, 1, 4), dtype=torch.float32, device="cuda")
    assert hm.blend_run_no_alloc(2, 2, 8000, 8000, 2, 1, 4, tiny, m, out) == HIP_ERROR_OUT_OF_MEMORY
    assert hm.blend_run_no_alloc(2, 2, 1, 1, 1, 1, 4, tiny, m, out) == 0
    torch.cuda.synchronize()
