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

// Head of one row of the network output z (48 floats): the softmax shift and
// scale of the mixture weights and the selection probability alpha are kept,
// each lobe (weight, unit mean, concentration) is decoded where it is used,
// so a renderer thread holds 4 values instead of the 49 of a decoded head.
struct GuideHead {
  const float* z;
  float mx, inv, alpha;
};

__device__ __forceinline__ void guide_head(const float* z, GuideHead& h) {
  float mx = z[0];
  for (int j = 1; j < kGLobes; j++) mx = fmaxf(mx, z[5 * j]);
  float s = 0.0f;
  for (int j = 0; j < kGLobes; j++) s += __expf(z[5 * j] - mx);
  h.z = z;
  h.mx = mx;
  h.inv = 1.0f / s;
  h.alpha = 0.05f + 0.9f / (1.0f + __expf(-z[40]));
}
__device__ __forceinline__ float guide_weight(const GuideHead& h, int j) { return __expf(h.z[5 * j] - h.mx) * h.inv; }
__device__ __forceinline__ void guide_lobe(const GuideHead& h, int j, float mu[3], float& kappa) {
  const float* z = h.z + 5 * j;
  float m0 = z[1], m1 = z[2], m2 = z[3];
  float l = fmaxf(sqrtf(m0 * m0 + m1 * m1 + m2 * m2), 1e-6f);
  mu[0] = m0 / l;
  mu[1] = m1 / l;
  mu[2] = m2 / l;
  float lk = fminf(8.0f, fmaxf(-3.0f, z[4]));
  kappa = __expf(lk);
}

__device__ __forceinline__ float guide_pdf(const GuideHead& h, Vec3 w) {
  float q = 0.0f;
  for (int j = 0; j < kGLobes; j++) {
    float mu[3], k;
    guide_lobe(h, j, mu, k);
    const float norm = k / (6.28318530717958647692f * (1.0f - __expf(-2.0f * k)));  // kappa / (2 pi (1 - e^-2kappa))
    float c = mu[0] * w.x + mu[1] * w.y + mu[2] * w.z;
    q += guide_weight(h, j) * norm * __expf(k * (c - 1.0f));
  }
  return q;
}

// Lobe pick by u0, then the numerically stable vMF sampler (W. Jakob,
// "Numerically stable sampling of the von Mises Fisher distribution on S^2", 2012).
__device__ __forceinline__ Vec3 guide_sample(const GuideHead& h, float u0, float u1, float u2) {
  int j = 0;
  float c = guide_weight(h, 0);
  while (j < kGLobes - 1 && u0 >= c) {
    j++;
    c += guide_weight(h, j);
  }
  float mu[3], k;
  guide_lobe(h, j, mu, k);
  float w = 1.0f + __logf(u1 + (1.0f - u1) * __expf(-2.0f * k)) / k;
  w = fminf(1.0f, fmaxf(-1.0f, w));
  float r = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  float sp, cp;
  __sincosf(6.28318530717958647692f * u2, &sp, &cp);
  Vec3 m = v3(mu[0], mu[1], mu[2]);
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

// Layout of the 96 values: (sin, cos) pairs, octave-major: position octave k
// of axis a at 6k + 2a, direction octave k of axis a at 72 + 6k + 2a.  out is
// 16-byte aligned (rows of 192 bytes); the row is written as 12 vector stores.
__device__ __forceinline__ void guide_encode(const GuideNet& N, Vec3 p, Vec3 d, uint16_t* out) {
  const float pc[3] = {p.x, p.y, p.z}, dc[3] = {d.x, d.y, d.z};
  uint32_t w[kGIn / 2];  // (sin, cos) bf16 pairs
  for (int a = 0; a < 3; a++) {
    float x = (pc[a] - N.lo[a]) / (N.hi[a] - N.lo[a]);
    x = fminf(1.0f, fmaxf(0.0f, x));
    float f = x;
    for (int k = 0; k < 12; k++) {
      float sn, cs;
      sincospif(f, &sn, &cs);
      w[3 * k + a] = (uint32_t)f2bf(sn) | ((uint32_t)f2bf(cs) << 16);
      f *= 2.0f;
    }
    f = dc[a];
    for (int k = 0; k < 4; k++) {
      float sn, cs;
      sincospif(f, &sn, &cs);
      w[36 + 3 * k + a] = (uint32_t)f2bf(sn) | ((uint32_t)f2bf(cs) << 16);
      f *= 2.0f;
    }
  }
  uint4* o4 = reinterpret_cast<uint4*>(out);
  for (int v = 0; v < kGIn / 8; v++) o4[v] = make_uint4(w[4 * v], w[4 * v + 1], w[4 * v + 2], w[4 * v + 3]);
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
