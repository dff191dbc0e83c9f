// Path-guiding network shared between the guide kernels and the renderer.
// No counterpart in the reference: a learned proposal (8-lobe von Mises-Fisher
// mixture + selection probability) for one-sample MIS with BSDF sampling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rx_dev.h"

namespace rx {

constexpr int kGIn = 96, kGHid = 64, kGOut = 48, kGLobes = 8;
constexpr int kGParams = kGIn * kGHid + 2 * kGHid * kGHid + kGHid * kGOut + 3 * kGHid + kGOut;  // 17,648

struct GuideNet {
  const uint16_t* Wt[4];  // bf16 [out][in]: B operand of the forward products
  const uint16_t* Wr[4];  // bf16 [in][out]: B operand of the backward products
  const float* b[4];
  float lo[3], hi[3];     // encoding bounds (scene box)
};

// Decoded head of one row: mixture weights, unit means, concentrations, alpha.
struct GuideHead {
  float pi[kGLobes];
  float mu[kGLobes][3];
  float kappa[kGLobes];
  float norm[kGLobes];  // kappa / (2 pi (1 - exp(-2 kappa)))
  float alpha;
};

__device__ __forceinline__ void guide_head(const float* z, GuideHead& h) {
  float mx = z[0];
  for (int j = 1; j < kGLobes; j++) mx = fmaxf(mx, z[5 * j]);
  float s = 0.0f;
  for (int j = 0; j < kGLobes; j++) {
    h.pi[j] = __expf(z[5 * j] - mx);
    s += h.pi[j];
  }
  const float inv = 1.0f / s;
  for (int j = 0; j < kGLobes; j++) {
    h.pi[j] *= inv;
    float m0 = z[5 * j + 1], m1 = z[5 * j + 2], m2 = z[5 * j + 3];
    float l = fmaxf(sqrtf(m0 * m0 + m1 * m1 + m2 * m2), 1e-6f);
    h.mu[j][0] = m0 / l;
    h.mu[j][1] = m1 / l;
    h.mu[j][2] = m2 / l;
    float lk = fminf(8.0f, fmaxf(-3.0f, z[5 * j + 4]));
    float k = __expf(lk);
    h.kappa[j] = k;
    h.norm[j] = k / (6.28318530717958647692f * (1.0f - __expf(-2.0f * k)));
  }
  h.alpha = 0.05f + 0.9f / (1.0f + __expf(-z[40]));
}

__device__ __forceinline__ float guide_pdf(const GuideHead& h, Vec3 w) {
  float q = 0.0f;
  for (int j = 0; j < kGLobes; j++) {
    float c = h.mu[j][0] * w.x + h.mu[j][1] * w.y + h.mu[j][2] * w.z;
    q += h.pi[j] * h.norm[j] * __expf(h.kappa[j] * (c - 1.0f));
  }
  return q;
}

// Lobe pick by u0, then the numerically stable vMF sampler (W. Jakob,
// "Numerically stable sampling of the von Mises Fisher distribution on S^2", 2012).
__device__ __forceinline__ Vec3 guide_sample(const GuideHead& h, float u0, float u1, float u2) {
  int j = 0;
  float c = h.pi[0];
  while (j < kGLobes - 1 && u0 >= c) {
    j++;
    c += h.pi[j];
  }
  const float k = h.kappa[j];
  float w = 1.0f + __logf(u1 + (1.0f - u1) * __expf(-2.0f * k)) / k;
  w = fminf(1.0f, fmaxf(-1.0f, w));
  float r = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  float sp, cp;
  __sincosf(6.28318530717958647692f * u2, &sp, &cp);
  Vec3 m = v3(h.mu[j][0], h.mu[j][1], h.mu[j][2]);
  Vec3 tmp = fabsf(m.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
  Vec3 t1 = normalize(cross(m, tmp));
  Vec3 t2 = cross(m, t1);
  return normalize(m * w + t1 * (r * cp) + t2 * (r * sp));
}

// Frequency encoding of one query into 96 bf16 values.
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void guide_encode(const GuideNet& N, Vec3 p, Vec3 d, uint16_t* out) {
  const float pc[3] = {p.x, p.y, p.z}, dc[3] = {d.x, d.y, d.z};
  for (int a = 0; a < 3; a++) {
    float x = (pc[a] - N.lo[a]) / (N.hi[a] - N.lo[a]);
    x = fminf(1.0f, fmaxf(0.0f, x));
    float f = x;
    for (int k = 0; k < 12; k++) {
      float s, c;
      sincospif(f, &s, &c);
      out[6 * k + 2 * a] = f2bf(s);
      out[6 * k + 2 * a + 1] = f2bf(c);
      f *= 2.0f;
    }
    f = dc[a];
    for (int k = 0; k < 4; k++) {
      float s, c;
      sincospif(f, &s, &c);
      out[72 + 6 * k + 2 * a] = f2bf(s);
      out[72 + 6 * k + 2 * a + 1] = f2bf(c);
      f *= 2.0f;
    }
  }
}

// Launchers (rx_guide.hip).  `rows` points at a device row count; the
// kernels grid-stride over 16-row tiles so any count up to `cap` works.
hipError_t guide_forward(const GuideNet& N, const uint16_t* X, const uint32_t* rows, uint32_t cap, uint16_t* A1,
                         uint16_t* A2, uint16_t* A3, float* Z4, int n_cu, hipStream_t s);

}  // namespace rx

#include "rx_context.h"

namespace rx {

// Network parameters (FP32 masters + Adam moments + BF16 copies) and the
// per-row buffers of inference and training.
struct GuideState {
  rx_guide_params prm{};
  DBuf params, adam_m, adam_v, grad, Wt[4], Wr[4];
  DBuf X, A1, A2, A3, Z4, dZ4, D1, D2, D3, partial[4], loss_part, rows, pos, dir, wdir, pb, wt;
  // one training pass's radiance samples, mini-batches are gathered from here
  DBuf spos, sdir, swdir, spb, swt, srows;
  uint32_t cap = 0, store_cap = 0;
  int step = 0;
  float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  GuideNet net() const {
    GuideNet n;
    for (int l = 0; l < 4; l++) {
      n.Wt[l] = Wt[l].as<uint16_t>();
      n.Wr[l] = Wr[l].as<uint16_t>();
    }
    const float* p = params.as<float>();
    const float* b = p + (kGIn * kGHid + 2 * kGHid * kGHid + kGHid * kGOut);
    n.b[0] = b;
    n.b[1] = b + kGHid;
    n.b[2] = b + 2 * kGHid;
    n.b[3] = b + 3 * kGHid;
    for (int a = 0; a < 3; a++) {
      n.lo[a] = lo[a];
      n.hi[a] = hi[a];
    }
    return n;
  }
};

int guide_reserve(rx_ctx* c, GuideState& G, uint32_t n);
int guide_reserve_store(rx_ctx* c, GuideState& G, uint32_t n);
int guide_train_rows(rx_ctx* c, uint32_t cap);

}  // namespace rx
