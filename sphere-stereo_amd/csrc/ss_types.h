// Internal types shared by host launch code and kernels.
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/sphere_stereo.h"

namespace ss {

// Double-sphere camera with the reference's torch-side scalars pre-rounded
// (utils.py:57-94): alpha**2, 2*alpha-1, 1-xi**2, 1-alpha and -w2 are Python
// doubles that enter fp32 tensor arithmetic as fp32 values.
struct TorchCam {
  float fs0, fs1, cs0, cs1;
  float xi, a, a2, am2_1, one_m_xi2, one_m_a, mw2;
};

inline TorchCam make_torch_cam(const ss_camera& c) {
  TorchCam t;
  double xi = c.xi, al = c.alpha;
  t.fs0 = c.fs[0];
  t.fs1 = c.fs[1];
  t.cs0 = c.cs[0];
  t.cs1 = c.cs[1];
  t.xi = (float)xi;
  t.a = (float)al;
  t.a2 = (float)(al * al);
  t.am2_1 = (float)(2 * al - 1);
  t.one_m_xi2 = (float)(1 - xi * xi);
  t.one_m_a = (float)(1 - al);
  double w1 = al > 0.5 ? (1 - al) / al : al / (1 - al);
  double w2 = (w1 + xi) / sqrt(2 * w1 * xi + xi * xi + 1);
  t.mw2 = (float)(-w2);
  return t;
}

// Candidates per lane V for a padded stride Np = 8 * V (lanes per pixel = 8).
inline int lanes_v(int N) {
  static const int table[] = {1, 2, 3, 4, 6, 8, 12, 16};
  for (int v : table)
    if (8 * v >= N) return v;
  return -1;
}

}  // namespace ss
