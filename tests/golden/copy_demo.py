"""Copies the reference's demo data (README.md:27-30: resources/ with the
four fisheye frames, their masks, the basalt calibration and the panorama
the reference produced, output/rgb_0.png) into tests/golden/ as fixtures.

Run once in a checkout next to the reference:
    python tests/golden/copy_demo.py /root/reference/resources
The frames and masks go to tests/golden/demo/ (laid out as a dataset for
python -m sphere_stereo.main); the expected panorama to
tests/golden/demo_expected_rgb_0.png, outside the dataset folder, so that a
run over tests/golden/demo/ cannot overwrite it.  Data only: no source file
of the reference is copied.
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(src):
    dst = os.path.join(HERE, "demo")
    os.makedirs(dst, exist_ok=True)
    shutil.copyfile(os.path.join(src, "calibration.json"), os.path.join(dst, "calibration.json"))
    for k in range(4):
        os.makedirs(os.path.join(dst, "cam%d" % k), exist_ok=True)
        for name in ("0.jpg", "mask.png"):
            shutil.copyfile(os.path.join(src, "cam%d" % k, name), os.path.join(dst, "cam%d" % k, name))
    shutil.copyfile(os.path.join(src, "output", "rgb_0.png"), os.path.join(HERE, "demo_expected_rgb_0.png"))


if __name__ == "__main__":
    main(sys.argv[1])
