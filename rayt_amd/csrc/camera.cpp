// Camera and 4x4 matrix math on the host (scene/camera.cpp:23-28,
// util/mat.cpp:6-194). The view-projection matrix is
// Perspective(fov_h, aspect, near, far = inf) * (RotY(-yaw) RotX(pitch) RotZ(-roll))^T * T(-position),
// inverted with the cofactor expansion of mat.cpp:20-107. fp32 throughout,
// no contraction; angles use the float libm functions.
#include <cmath>
#include <cstring>

#include "rt_internal.h"

namespace rayt {

struct Mat4 {
  float m[16];

  static Mat4 Zero() {
    Mat4 r;
    for (float& v : r.m) v = 0.0f;
    return r;
  }
  static Mat4 Identity() {
    Mat4 r = Zero();
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
  }
  Mat4 operator*(const Mat4& b) const {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        float t = 0.0f;
        for (int k = 0; k < 4; ++k) t += m[i * 4 + k] * b.m[k * 4 + j];
        r.m[i * 4 + j] = t;
      }
    return r;
  }
  Mat4 Transposed() const {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) r.m[j * 4 + i] = m[i * 4 + j];
    return r;
  }
  // Pairs-of-products cofactor inverse; each destination element sums three
  // products, subtracts the complementary three, then scales by 1/det.
  Mat4 Inverse() const {
    float s[16];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) s[j * 4 + i] = m[i * 4 + j];
    float p[12];
    Mat4 r;
    float* d = r.m;
    p[0] = s[10] * s[15]; p[1] = s[11] * s[14]; p[2] = s[9] * s[15]; p[3] = s[11] * s[13];
    p[4] = s[9] * s[14]; p[5] = s[10] * s[13]; p[6] = s[8] * s[15]; p[7] = s[11] * s[12];
    p[8] = s[8] * s[14]; p[9] = s[10] * s[12]; p[10] = s[8] * s[13]; p[11] = s[9] * s[12];
    auto cof = [&](int a0, int b0, int a1, int b1, int a2, int b2, int c0, int e0, int c1, int e1, int c2, int e2) {
      float v = p[a0] * s[b0] + p[a1] * s[b1] + p[a2] * s[b2];
      v -= p[c0] * s[e0] + p[c1] * s[e1] + p[c2] * s[e2];
      return v;
    };
    d[0] = cof(0, 5, 3, 6, 4, 7, 1, 5, 2, 6, 5, 7);
    d[1] = cof(1, 4, 6, 6, 9, 7, 0, 4, 7, 6, 8, 7);
    d[2] = cof(2, 4, 7, 5, 10, 7, 3, 4, 6, 5, 11, 7);
    d[3] = cof(5, 4, 8, 5, 11, 6, 4, 4, 9, 5, 10, 6);
    d[4] = cof(1, 1, 2, 2, 5, 3, 0, 1, 3, 2, 4, 3);
    d[5] = cof(0, 0, 7, 2, 8, 3, 1, 0, 6, 2, 9, 3);
    d[6] = cof(3, 0, 6, 1, 11, 3, 2, 0, 7, 1, 10, 3);
    d[7] = cof(4, 0, 9, 1, 10, 2, 5, 0, 8, 1, 11, 2);
    p[0] = s[2] * s[7]; p[1] = s[3] * s[6]; p[2] = s[1] * s[7]; p[3] = s[3] * s[5];
    p[4] = s[1] * s[6]; p[5] = s[2] * s[5]; p[6] = s[0] * s[7]; p[7] = s[3] * s[4];
    p[8] = s[0] * s[6]; p[9] = s[2] * s[4]; p[10] = s[0] * s[5]; p[11] = s[1] * s[4];
    d[8] = cof(0, 13, 3, 14, 4, 15, 1, 13, 2, 14, 5, 15);
    d[9] = cof(1, 12, 6, 14, 9, 15, 0, 12, 7, 14, 8, 15);
    d[10] = cof(2, 12, 7, 13, 10, 15, 3, 12, 6, 13, 11, 15);
    d[11] = cof(5, 12, 8, 13, 11, 14, 4, 12, 9, 13, 10, 14);
    d[12] = cof(2, 10, 5, 11, 1, 9, 4, 11, 0, 9, 3, 10);
    d[13] = cof(8, 11, 0, 8, 7, 10, 6, 10, 9, 11, 1, 8);
    d[14] = cof(6, 9, 11, 11, 3, 8, 10, 11, 2, 8, 7, 9);
    d[15] = cof(10, 10, 4, 8, 9, 9, 8, 9, 11, 10, 5, 8);
    float det = s[0] * d[0] + s[1] * d[1] + s[2] * d[2] + s[3] * d[3];
    det = 1.0f / det;
    for (float& v : r.m) v *= det;
    return r;
  }
};

static const float kPi = 3.1415926535897932384626433832795f;
static const float kDegToRad = kPi / 180.0f;

static Mat4 Rotation(int axis, float deg) {
  const float c = std::cos(deg * kDegToRad), s = std::sin(deg * kDegToRad);
  Mat4 r = Mat4::Identity();
  if (axis == 0) {
    r.m[5] = c; r.m[6] = -s; r.m[9] = s; r.m[10] = c;
  } else if (axis == 1) {
    r.m[0] = c; r.m[2] = s; r.m[8] = -s; r.m[10] = c;
  } else {
    r.m[0] = c; r.m[1] = -s; r.m[4] = s; r.m[5] = c;
  }
  return r;
}

static Mat4 Perspective(float fov, float aspect, float near_dist) {
  Mat4 r = Mat4::Zero();
  const float f = 1.0f / std::tan(fov * (kDegToRad / 2));
  r.m[0] = f;
  r.m[5] = f * aspect;
  r.m[14] = -1.0f;
  r.m[10] = -1.0f;
  r.m[11] = -2 * near_dist;
  return r;
}

}  // namespace rayt

extern "C" int rt_camera_view_proj_inv(float yaw, float pitch, float roll, const float position[3], float fov_deg,
                                       float near_clip, float aspect, float out16[16]) {
  RT_REQUIRE(position && out16, RT_ERR_INVALID, "null argument");
  RT_REQUIRE(fov_deg > 0 && fov_deg < 180 && near_clip > 0 && aspect > 0, RT_ERR_RANGE, "bad camera parameters");
  using rayt::Mat4;
  Mat4 t = Mat4::Identity();
  t.m[3] = -position[0];
  t.m[7] = -position[1];
  t.m[11] = -position[2];
  const Mat4 rot = rayt::Rotation(1, -yaw) * rayt::Rotation(0, pitch) * rayt::Rotation(2, -roll);
  const Mat4 vp = rayt::Perspective(fov_deg, aspect, near_clip) * rot.Transposed() * t;
  const Mat4 inv = vp.Inverse();
  memcpy(out16, inv.m, sizeof(inv.m));
  return RT_OK;
}
