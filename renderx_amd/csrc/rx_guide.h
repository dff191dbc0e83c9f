// Path-guiding network shared between the guide kernels and the renderer.
// No counterpart in the reference: a learned proposal (8-lobe von Mises-Fisher
// mixture + selection probability) for one-sample MIS with BSDF sampling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rx_dev.h"

namespace rx {

constexpr int kGIn = 96, kGHid = 64, kGOut = 48, kGLobes = 8;
// Selection probability of the learned proposal: alpha = kAlphaMin +
// kAlphaSpan sigmoid(z[40]) in [0.05, 0.5].  Keeping at least half of every
// guided vertex's samples on the BSDF / phase function bounds the vertex's
// weight f / p_mix by twice the plain strategy's, so throughputs cannot grow
// geometrically along a path whose directions the mixture misses.
#ifndef RX_GUIDE_ALPHA_MAX
#define RX_GUIDE_ALPHA_MAX 0.5f
#endif
constexpr float kAlphaMin = 0.05f, kAlphaSpan = RX_GUIDE_ALPHA_MAX - kAlphaMin;
constexpr int kGParams = kGIn * kGHid + 2 * kGHid * kGHid + kGHid * kGOut + 3 * kGHid + kGOut;  // 17,648

struct GuideNet {
  const uint16_t* Wt[4];  // bf16 [out][in]: B operand of the forward products
  const uint16_t* Wr[4];  // bf16 [in][out]: B operand of the backward products
  const float* b[4];
  float lo[3], hi[3];     // encoding bounds (scene box)
};

// Decoded network output of one renderer query, kGHead bytes, all FP32:
//   float4 lobe[8]  unit mean mu_j and concentration kappa_j
//   float  cw[8]    w_j kappa_j / (2 pi (1 - e^{-2 kappa_j})): weight times vMF normaliser
//   float  cdf[8]   cumulative weights, cdf[7] = 1; w_j = cdf[j] - cdf[j-1]
//   float  alpha    selection probability of the mixture (then 3 floats of padding)
// The decode computes the weights from the stored cumulative table, so the
// sampler (lobe j where u0 < cdf[j] first holds) and the density read the
// same numbers: the density is exactly the sampler's up to FP32 rounding.
typedef _Float16 half_t;
constexpr int kGHead = 8 * 16 + 8 * 4 + 8 * 4 + 16;  // 208 bytes per row
constexpr int kGHeadCw = 128, kGHeadCdf = 160, kGHeadAlpha = 192;

struct GuideHead {
  const float4* lobe;
  const float* cw;
  const float* cdf;
  float alpha;
};

__device__ __forceinline__ void guide_head(const unsigned char* row, GuideHead& g) {
  g.lobe = reinterpret_cast<const float4*>(row);
  g.cw = reinterpret_cast<const float*>(row + kGHeadCw);
  g.cdf = reinterpret_cast<const float*>(row + kGHeadCdf);
  g.alpha = *reinterpret_cast<const float*>(row + kGHeadAlpha);
}

__device__ __forceinline__ float guide_pdf(const GuideHead& h, Vec3 w) {
  float q = 0.0f;
#pragma unroll
  for (int j = 0; j < kGLobes; j++) {
    const float4 L = h.lobe[j];
    const float c = L.x * w.x + L.y * w.y + L.z * w.z;
    q += h.cw[j] * __expf(L.w * (c - 1.0f));
  }
  return q;
}

// Lobe pick by u0, then the numerically stable vMF sampler (W. Jakob,
// "Numerically stable sampling of the von Mises Fisher distribution on S^2", 2012).
__device__ __forceinline__ Vec3 guide_sample(const GuideHead& h, float u0, float u1, float u2) {
  int j = 0;
  while (j < kGLobes - 1 && u0 >= h.cdf[j]) j++;
  const float4 L = h.lobe[j];
  const Vec3 m = v3(L.x, L.y, L.z);
  const float k = L.w;
  float w = 1.0f + __logf(u1 + (1.0f - u1) * __expf(-2.0f * k)) * __builtin_amdgcn_rcpf(k);
  w = fminf(1.0f, fmaxf(-1.0f, w));
  float r = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  float sp, cp;
  __sincosf(6.28318530717958647692f * u2, &sp, &cp);
  Vec3 tmp = fabsf(m.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
  Vec3 t1 = normalize(cross(m, tmp));
  Vec3 t2 = cross(m, t1);
  return normalize(m * w + t1 * (r * cp) + t2 * (r * sp));
}

// FP32 -> BF16, round to nearest even: gfx950's v_cvt_pk_bf16_f32, two
// values per instruction for bf2_pack.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// (bf16(lo) | bf16(hi) << 16)
__device__ __forceinline__ uint32_t bf2_pack(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){lo, hi}, bf16x2_t));
}
// Frequency encoding of one query into 96 bf16 values.
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
      w[3 * k + a] = bf2_pack(sn, cs);
      f *= 2.0f;
    }
    f = dc[a];
    for (int k = 0; k < 4; k++) {
      float sn, cs;
      sincospif(f, &sn, &cs);
      w[36 + 3 * k + a] = bf2_pack(sn, cs);
      f *= 2.0f;
    }
  }
  uint4* o4 = reinterpret_cast<uint4*>(out);
  for (int v = 0; v < kGIn / 8; v++) o4[v] = make_uint4(w[4 * v], w[4 * v + 1], w[4 * v + 2], w[4 * v + 3]);
}

// One (sin, cos) pair of the encoding: position octave k of axis a
// (pair 3k + a, k < 12) or direction octave k (pair 36 + 3k + a, k < 4),
// the same values guide_encode writes.
__device__ __forceinline__ uint32_t guide_pair(const GuideNet& N, Vec3 p, Vec3 d, int pair) {
  const float pc[3] = {p.x, p.y, p.z}, dc[3] = {d.x, d.y, d.z};
  float f;
  if (pair < 36) {
    const int k = pair / 3, a = pair % 3;
    float x = (pc[a] - N.lo[a]) / (N.hi[a] - N.lo[a]);
    x = fminf(1.0f, fmaxf(0.0f, x));
    f = ldexpf(x, k);  // x doubled k times, exactly
  } else {
    const int k = (pair - 36) / 3, a = (pair - 36) % 3;
    f = ldexpf(dc[a], k);
  }
  float sn, cs;
  sincospif(f, &sn, &cs);
  return bf2_pack(sn, cs);
}

// ------------------------------------------------------------------ MFMA tiles
// Every product of the network has one dimension of 16 rows per wavefront
// (a tile of queries).  Operands are kept as the fragments of K = 16 steps
// (the layout of v_mfma_f32_16x16x16_bf16):
//   A fragment: lane l holds A[l%16][4(l/16) .. 4(l/16)+3]
//   B fragment: lane l holds B[4(l/16) .. 4(l/16)+3][l%16]
//   D fragment: lane l holds D[4(l/16)+i][l%16], i = 0..3
// and two consecutive K = 16 steps run as one v_mfma_f32_16x16x32_bf16,
// gfx950's full-rate BF16 product (the K = 16 form issues at about half the
// rate): that instruction's lane l holds A[l%16][8(l/16) + j] and
// B[8(l/16) + j][l%16], j = 0..7, so lane l's K = 16 fragments of steps 2s
// and 2s + 1 placed side by side give both operands the same permutation of
// k within the 32-wide step, and the product is the sum of the two steps.
typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kLayerIn[4] = {kGIn, kGHid, kGHid, kGHid};
constexpr int kLayerOut[4] = {kGHid, kGHid, kGHid, kGOut};

// Four consecutive bf16 values (8-byte aligned) as a fragment, the loaded
// dwords reinterpreted in place.
__device__ __forceinline__ v4s ld4(const uint16_t* p) {
  return __builtin_bit_cast(v4s, *reinterpret_cast<const uint2*>(p));
}
// Two K = 16 steps (fragments a0, b0 and a1, b1) in one K = 32 product; a
// product over an odd number of K = 16 steps pads the last with zero
// fragments.  Every product of the network uses this one instruction: an
// accumulator passed from a 16x16x32 product to a 16x16x16 one came out wrong
// in the build tested (the compiler spaced the dependent pair by one
// instruction), so the two forms are not mixed.
__device__ __forceinline__ v4f mfma2(v4s a0, v4s a1, v4s b0, v4s b1, v4f c) {
  const v8s a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
  const v8s b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                 0);
}

// The forward weights in LDS, [out][in] per layer with rows padded by 8 bf16
// so that the 16 rows of a fragment load fall in different banks.
constexpr int kWPad = 8;
constexpr int kWOff[4] = {0, kGHid * (kGIn + kWPad), kGHid * (kGIn + kWPad) + kGHid * (kGHid + kWPad),
                          kGHid * (kGIn + kWPad) + 2 * kGHid * (kGHid + kWPad)};
constexpr int kWLds = kWOff[3] + kGOut * (kGHid + kWPad);  // bf16 values

// Block-cooperative copy of the forward weights into LDS (wts: kWLds values).
__device__ __forceinline__ void stage_weights(const GuideNet& N, uint16_t* wts) {
  for (int L = 0; L < 4; L++) {
    const int K = kLayerIn[L], NO = kLayerOut[L];
    for (int e = threadIdx.x; e < NO * K / 4; e += blockDim.x) {  // 4 bf16 per step
      const int r = (4 * e) / K, col = (4 * e) % K;
      *reinterpret_cast<uint2*>(&wts[kWOff[L] + r * (K + kWPad) + col]) =
          *reinterpret_cast<const uint2*>(N.Wt[L] + (size_t)r * K + col);
    }
  }
}

// Output order of the decoding inference: staged output row 16 mt + 4 g + i
// (the D fragment of output tile mt that lane group g = lane / 16 holds, its
// slot s = 4 mt + i) is network output s < 10 ? 10 g + s : 40 + 2 g + s - 10.
// Lane group g then holds lobes 2g and 2g + 1 whole (weight logit, mean,
// log kappa in slots 0-4 and 5-9), group 0 the selection logit (output 40)
// in slot 10, and the unused outputs 41-47 fill the remaining slots.
__device__ __forceinline__ int decode_order(int row) {
  const int g = (row >> 2) & 3, s = 4 * (row >> 4) + (row & 3);
  return s < 10 ? 10 * g + s : 40 + 2 * g + s - 10;
}

// The four layers' biases in LDS, b1 | b2 | b3 | b4 (kBiasCount floats),
// b4 in decode_order with out_order.
constexpr int kBiasCount = 3 * kGHid + kGOut;
__device__ __forceinline__ void stage_bias(const GuideNet& N, float* bias, bool out_order) {
  for (int e = threadIdx.x; e < kBiasCount; e += blockDim.x) {
    const int L = e / kGHid, o = e % kGHid;
    bias[e] = N.b[L][L == 3 && out_order ? decode_order(o) : o];
  }
}

// Inference order of the encoding (k_guide_fwd): at k-step ks the lanes of
// group lg (lane / 16) of a 16-row tile supply elements 2 ks and 2 ks + 1 of
// their group's chain of 12 pairs: groups 0, 1, 2 the 12 position octaves of
// axis lg, group 3 the 4 direction octaves of each axis in turn.  Along a
// chain each octave follows from the previous one by the double-angle
// formulas, so a lane evaluates sincospi once per axis; the first layer's
// weight columns are staged in this order.
__device__ __forceinline__ int enc_chain_pair(int lg, int m) {
  return lg < 3 ? 3 * m + lg : 36 + 3 * (m & 3) + (m >> 2);
}
// stage_weights with the first layer's input columns in the chain order:
// staged column 16 ks + 4 lg + 2 u + v holds feature 2 enc_chain_pair(lg, 2 ks + u) + v;
// with out_order the last layer's rows are in decode_order.
__device__ __forceinline__ void stage_weights_chain(const GuideNet& N, uint16_t* wts, bool out_order) {
  for (int e = threadIdx.x; e < kGHid * kGIn; e += blockDim.x) {
    const int r = e / kGIn, c = e % kGIn;
    const int ks = c >> 4, lg = (c >> 2) & 3, u = (c >> 1) & 1, v = c & 1;
    wts[kWOff[0] + r * (kGIn + kWPad) + c] = N.Wt[0][(size_t)r * kGIn + 2 * enc_chain_pair(lg, 2 * ks + u) + v];
  }
  for (int L = 1; L < 4; L++) {
    const int K = kLayerIn[L], NO = kLayerOut[L];
    for (int e = threadIdx.x; e < NO * K / 4; e += blockDim.x) {  // 4 bf16 per step
      const int r = (4 * e) / K, col = (4 * e) % K;
      const int src = L == 3 && out_order ? decode_order(r) : r;
      *reinterpret_cast<uint2*>(&wts[kWOff[L] + r * (K + kWPad) + col]) =
          *reinterpret_cast<const uint2*>(N.Wt[L] + (size_t)src * K + col);
    }
  }
}

// Launchers (rx_guide.hip).  `rows` points at a device row count; the
// kernels grid-stride over 16-row tiles so any count up to `cap` works.
// Encode + forward of `rows` queries (pos[r * stride], dir[r * stride]):
// raw outputs into Z4 (48 fp32 per row), or with head != NULL the decoded
// mixtures (kGHead bytes per row).
hipError_t guide_infer(const GuideNet& N, const float4* pos, const float4* dir, uint32_t stride, const uint32_t* rows,
                       uint32_t cap, float* Z4, unsigned char* head, int n_cu, hipStream_t s);


}  // namespace rx

#include "rx_context.h"

namespace rx {

// Network parameters (FP32 masters + Adam moments + BF16 copies) and the
// per-row buffers of inference and training.
struct GuideState {
  rx_guide_params prm{};
  DBuf params, adam_m, adam_v, grad, Wt[4], Wr[4];
  DBuf Z4, tpart, loss_part, rows, pos, dir, wdir, pb, wt;  // rx_guide_eval outputs, training partials, step rows
  // one training pass's radiance samples, mini-batches are gathered from here
  DBuf spos, sdir, swdir, spb, swt, srows;
  uint32_t cap = 0, store_cap = 0, train_blocks = 1;
  int step = 0;
  uint32_t decay_steps = 0;  // Adam steps of the online warm-up (RX_GUIDE_LR_DECAY)
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
