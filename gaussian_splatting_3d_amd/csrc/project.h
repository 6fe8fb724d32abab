// Per-Gaussian projection math shared by the single-view and multi-view
// preprocessing kernels and their adjoint (gs/renderer.py:365-419,
// gs/culling.py:8-37, culling.h:11-20, kernels.h:156-170).  Every value that
// feeds a decision or a key follows the reference's operation order.
#pragma once

#include "gsmi_device.h"

namespace gsmi {

struct Planes {
  float n[18];
  float p[18];
};

// kernels.h:156-170 intersect_sphere_frustum.
__device__ __forceinline__ bool in_frustum(float mx, float my, float mz, float r, const float* __restrict__ nrm,
                                           const float* __restrict__ pts) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const float dx = mx - pts[3 * k], dy = my - pts[3 * k + 1], dz = mz - pts[3 * k + 2];
    const float d = dx * nrm[3 * k] + dy * nrm[3 * k + 1] + dz * nrm[3 * k + 2];
    if (!(d > -r)) return false;
  }
  return true;
}

// utils/transforms.py:31-45 (kornia WXYZ): normalise, then the unit
// quaternion rotation matrix, row-major.
__device__ __forceinline__ void quat_to_rot(float4 q, float* R) {
  const float nrm = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  const float den = fmaxf(nrm, 1e-12f);
  const float w = q.x / den, x = q.y / den, y = q.z / den, z = q.w / den;
  const float tx = 2.0f * x, ty = 2.0f * y, tz = 2.0f * z;
  const float twx = tx * w, twy = ty * w, twz = tz * w;
  const float txx = tx * x, txy = ty * x, txz = tz * x;
  const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
  R[0] = 1.0f - (tyy + tzz);
  R[1] = txy - twz;
  R[2] = txz + twy;
  R[3] = txy + twz;
  R[4] = 1.0f - (txx + tzz);
  R[5] = tyz - twx;
  R[6] = txz - twy;
  R[7] = tyz + twx;
  R[8] = 1.0f - (txx + tyy);
}

struct Proj {
  float mx, my;  // mean2d
  float c[4];    // cov2d row-major
  float depth;
};

// gs/renderer.py:380-419 with c2w = [Rc | t]: u = Rc^T (mu - t),
// Sigma = (R(q) diag s)(R(q) diag s)^T, A = J Rc^T (J detached),
// cov2d = (A Sigma A^T)[:2,:2], depth = u_z, mean2d = u_xy / depth.
__device__ __forceinline__ Proj project_one(float m0, float m1, float m2, float4 q, float s0, float s1, float s2,
                                            const float* __restrict__ cw) {
  float Wm[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Wm[3 * i + j] = cw[4 * j + i];
  const float d0 = m0 + (-cw[3]), d1 = m1 + (-cw[7]), d2 = m2 + (-cw[11]);
  float u[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) u[i] = Wm[3 * i + 0] * d0 + Wm[3 * i + 1] * d1 + Wm[3 * i + 2] * d2;
  float R[9], M[9], S[9];
  quat_to_rot(q, R);
  const float s[3] = {s0, s1, s2};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[3 * i + j] = s[j] * R[3 * i + j];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      S[3 * a + b] = M[3 * a + 0] * M[3 * b + 0] + M[3 * a + 1] * M[3 * b + 1] + M[3 * a + 2] * M[3 * b + 2];
  const float x = u[0], y = u[1], z = u[2];
  const float J00 = 1.0f / z, J02 = -x / z / z;
  const float J11 = 1.0f / z, J12 = -y / z / z;
  float A[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A[k] = J00 * Wm[0 * 3 + k] + J02 * Wm[2 * 3 + k];
    A[3 + k] = J11 * Wm[1 * 3 + k] + J12 * Wm[2 * 3 + k];
  }
  float T1[6];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      T1[3 * i + b] = A[3 * i + 0] * S[0 * 3 + b] + A[3 * i + 1] * S[1 * 3 + b] + A[3 * i + 2] * S[2 * 3 + b];
  Proj p;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
      p.c[2 * i + j] = T1[3 * i + 0] * A[3 * j + 0] + T1[3 * i + 1] * A[3 * j + 1] + T1[3 * i + 2] * A[3 * j + 2];
  p.depth = z;
  p.mx = x / z;
  p.my = y / z;
  return p;
}

// gs/culling.py:16-31 with utils/camera.py:296-301 for one axis.
__device__ __forceinline__ void aabb_axis(float m, float cii, float D, float f, float c, int lim, int tile,
                                          int* lo, int* hi) {
  const float e = sqrtf(D * cii);
  const float pl = (m - e) * f;
  const float ph = (m + e) * f;
  int il = f2i_sat(pl + c);
  int ih = f2i_sat(ph + c);
  il = il < 0 ? 0 : (il > lim ? lim : il);
  ih = ih < 0 ? 0 : (ih > lim ? lim : ih);
  *lo = il / tile;
  *hi = ih / tile;
}

}  // namespace gsmi
