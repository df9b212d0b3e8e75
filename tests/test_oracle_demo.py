"""The oracle on the reference's README demo, against the reference's own
output.

README.md:27-30 runs ``python python/main.py --dataset_path resources
--references_indices 2 3`` with main.py's defaults (1024^2 matching, 32
candidates, sigma_i 10, sigma_s 25, 1216^2 stitch input, 2048 x 1024
panorama); resources/output/rgb_0.png is what the reference wrote.  The
demo's frames, masks, calibration and that panorama are committed under
tests/golden/.  The oracle cannot match it bit for bit: the reference ran
CUDA's __expf, NVRTC's default FMA contraction, racy reprojection and
in-place inpainting kernels, and cv2's JPEG decoder and uint8 mask resize.
What it pins is the whole composition at full size: a sign slip in the
merge's colour reprojection falls to ~20 dB, an ISB sigma-scaling slip to
~38 dB; the oracle gives 47.5 dB with 53 % of the pixels exact.
"""
import numpy as np

from helpers import DEMO_EXPECTED, demo_frame, psnr
from oracle import Oracle
from pipeline import OraclePipeline


def test_demo_panorama_matches_the_reference_output():
    from sphere_stereo.main import read_image

    o = Oracle("f32")
    calibs, kw, to_match, to_stitch = demo_frame(o)
    rgb, dist = OraclePipeline(calibs, **kw).estimate_RGBD_panorama(to_match, to_stitch)
    want = read_image(DEMO_EXPECTED)
    assert rgb.shape == want.shape == (1024, 2048, 3)
    p = psnr(rgb, want)
    exact = np.mean(np.all(rgb == want, axis=-1))
    near = np.mean(np.abs(rgb.astype(int) - want.astype(int)) <= 8)
    assert p >= 45.0, "PSNR %.2f dB against the reference's panorama" % p
    assert exact >= 0.50, "%.1f %% of the pixels exact" % (100 * exact)
    assert near >= 0.995, "%.2f %% of the values within 8 levels" % (100 * near)
    # the distance panorama: finite where the blend weights are, within the candidate range
    ok = np.isfinite(dist)
    assert ok.mean() > 0.85
    assert dist[ok].min() >= 0.55 * (1 - 1e-6) and dist[ok].max() <= 100 * (1 + 1e-6)
