"""Writes known_answers.json: the input/expected vectors of hm-cupano's own
blend unit tests, transcribed from the test sources (no reference code is
imported or run).  Each case names the test it comes from.

Run: python tests/golden/make_known_answers.py
"""
import json
import os

CASES = [
    # src/cuda/cudaBlendN_test.cpp:27-71
    dict(name="blendN_alpha_zero_skips_contribution", source="src/cuda/cudaBlendN_test.cpp:27-71",
         kind="N", w=1, h=1, c=4, batch=1, levels=1,
         images=[[10, 20, 30, 255], [100, 110, 120, 0], [90, 100, 110, 255]],
         mask=[0.1, 0.8, 0.1], expected=[50, 60, 70, 255], tol=1e-5),
    # src/cuda/cudaBlendN_test.cpp:142-192
    dict(name="blendN_mask_selects_transparent_then_fallback", source="src/cuda/cudaBlendN_test.cpp:142-192",
         kind="N", w=1, h=1, c=4, batch=1, levels=1,
         images=[[255, 0, 0, 255], [0, 255, 0, 0], [0, 0, 255, 128]],
         mask=[0.0, 1.0, 0.0], expected=[255, 0, 0, 255], tol=1e-5),
    # src/cuda/cudaBlend3_test.cpp:66 (three one-hot masks, RGB)
    dict(name="blend3_single_pixel_identity_mask_0", source="src/cuda/cudaBlend3_test.cpp:66-150",
         kind="3", w=1, h=1, c=3, batch=1, levels=1,
         images=[[10, 20, 30], [40, 50, 60], [70, 80, 90]], mask=[1, 0, 0], expected=[10, 20, 30], tol=1e-5),
    dict(name="blend3_single_pixel_identity_mask_1", source="src/cuda/cudaBlend3_test.cpp:66-150",
         kind="3", w=1, h=1, c=3, batch=1, levels=1,
         images=[[10, 20, 30], [40, 50, 60], [70, 80, 90]], mask=[0, 1, 0], expected=[40, 50, 60], tol=1e-5),
    dict(name="blend3_single_pixel_identity_mask_2", source="src/cuda/cudaBlend3_test.cpp:66-150",
         kind="3", w=1, h=1, c=3, batch=1, levels=1,
         images=[[10, 20, 30], [40, 50, 60], [70, 80, 90]], mask=[0, 0, 1], expected=[70, 80, 90], tol=1e-5),
    # src/cuda/cudaBlend3_test.cpp:198-250
    dict(name="blend3_single_pixel_rgba", source="src/cuda/cudaBlend3_test.cpp:198-250",
         kind="3", w=1, h=1, c=4, batch=1, levels=1,
         images=[[10, 20, 30, 40], [50, 60, 70, 80], [90, 100, 110, 120]],
         mask=[0.2, 0.3, 0.5], expected=[62, 72, 82, 120], tol=1e-5),
    # src/cuda/cudaBlend3_test.cpp:255-290
    dict(name="blend3_alpha_zero_skips_contribution", source="src/cuda/cudaBlend3_test.cpp:255-290",
         kind="3", w=1, h=1, c=4, batch=1, levels=1,
         images=[[10, 20, 30, 255], [100, 110, 120, 0], [90, 100, 110, 255]],
         mask=[0.1, 0.8, 0.1], expected=[50, 60, 70, 255], tol=1e-5),
    # src/cuda/cudaBlend3_test.cpp:293-330
    dict(name="blend3_mask_selects_transparent_then_fallback", source="src/cuda/cudaBlend3_test.cpp:293-330",
         kind="3", w=1, h=1, c=4, batch=1, levels=1,
         images=[[255, 0, 0, 255], [0, 255, 0, 0], [0, 0, 255, 128]],
         mask=[0, 1, 0], expected=[255, 0, 0, 255], tol=1e-5),
]

# Cases whose inputs are generated by a rule rather than listed.
RULE_CASES = [
    # src/cuda/cudaBlendN_test.cpp:74-139: constant images, weights .25/.75, 8x8, B=2, L=4
    dict(name="blendN_multilevel_constant_matches_weights", source="src/cuda/cudaBlendN_test.cpp:74-139",
         kind="N", w=8, h=8, c=4, batch=2, levels=4,
         constants=[[10, 20, 30, 255], [110, 120, 130, 255]], weights=[0.25, 0.75],
         expected=[85, 95, 105, 255], tol=1e-5),
    # src/cuda/cudaBlend3_test.cpp:155-195: 2x2 ones/twos/threes, uniform 1/3 mask
    dict(name="blend3_two_by_two_uniform_mask", source="src/cuda/cudaBlend3_test.cpp:155-195",
         kind="3", w=2, h=2, c=3, batch=1, levels=1,
         constants=[[1, 1, 1], [2, 2, 2], [3, 3, 3]], weights=[1 / 3, 1 / 3, 1 / 3],
         expected=[2, 2, 2], tol=1e-5),
]

# src/cuda/cudaBlend3_test.cpp:333-380: 4x4, L=2, red/green/blue, mask top/bottom -> no alpha holes
NO_HOLES = dict(name="blend3_multilevel_overlap_has_no_alpha_holes", source="src/cuda/cudaBlend3_test.cpp:333-380",
                w=4, h=4, c=4, levels=2, min_alpha=250.0, min_rgb_sum=50.0)

# tests/python/test_pytorch_pano.py:192-201: Laplacian alpha renormalisation
LAPLACIAN_ALPHA = dict(name="laplacian_renormalizes_fractional_alpha_weights",
                       source="tests/python/test_pytorch_pano.py:192-201",
                       high_shape=[1, 4, 4, 4], low_shape=[1, 2, 2, 4], low_pixel=[1, 1], low_value=[100, 110, 120, 255],
                       probe=[1, 1], expected_rgb=[-100, -110, -120], expected_alpha=0.0, tol=1e-5)

if __name__ == "__main__":
    out = dict(cases=CASES, rule_cases=RULE_CASES, no_holes=NO_HOLES, laplacian_alpha=LAPLACIAN_ALPHA)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)
