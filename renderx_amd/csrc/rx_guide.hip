// Path-guiding network on gfx950 matrix cores: the renderer's inference
// (encode + forward + decode) and the fused training step (forward, loss,
// backward and weight gradients in one kernel, then a deterministic
// block-order reduction with the FP32 Adam update), all products on
// v_mfma_f32_16x16x16_bf16 with FP32 accumulation (fragment layout in
// rx_guide.h).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "rx_context.h"
#include "rx_guide.h"

namespace rx {


// ------------------------------------------------------------------ inference
// Encode + forward for queries given as (position, direction to the viewer),
// one 16-row tile per wave at a time, waves independent of each other once
// the block has staged the forward weights in LDS.  Each layer computes the
// transposed product H_L^T = W_L H_{L-1}^T (out x rows) with the weights as
// the A operand; the D fragment of its output tile mt (lane l holds
// H^T[16 mt + 4(l/16) + i][l%16]) is then exactly the B fragment of the next
// layer's k-step mt, so the activations never leave the registers, and each
// lane computes the encoding features of its own B fragments (row l%16,
// features 16 ks + 4(l/16) .. +3).  Outputs: the raw fp32 outputs (Z4, 48 per
// row) or, with DECODE, the decoded mixture as kGHead bytes per row (the
// renderer's form, GuideHead), decoded in the registers of the last layer's
// D fragments with its outputs staged in decode_order.  Query row r is
// pos[r * stride], dir[r * stride].
// 16 waves share one copy of the weights (39.6 KB of LDS with the biases);
// two blocks per CU, 8 waves per SIMD at 64 VGPRs
constexpr int kFwdBlock = 1024, kFwdBlocksPerCu = 2;
constexpr int kFwdWaves = kFwdBlock / 64;

// One layer: acc[mt] = W[16 mt .. , :] * in + bias for the layer's MT output
// tiles, the bias loaded as the products' starting accumulator (lane l's four
// outputs 16 mt + 4(l/16) .. +3 are consecutive, 16-byte aligned floats).
template <int KS, int MT>
__device__ __forceinline__ void fwd_layer(const uint16_t* w, int ldw, const float* bias, const v4s* in, v4f* acc,
                                          int lr, int lg) {
  static_assert(KS % 2 == 0, "layer inputs in whole 32-wide steps");
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
    const float4 b4 = *reinterpret_cast<const float4*>(bias + mt * 16 + 4 * lg);
    v4f a = v4f{b4.x, b4.y, b4.z, b4.w};
    const uint16_t* wr = w + (mt * 16 + lr) * ldw + 4 * lg;
#pragma unroll
    for (int ks = 0; ks < KS; ks += 2) a = mfma2(ld4(wr + ks * 16), ld4(wr + ks * 16 + 16), in[ks], in[ks + 1], a);
    acc[mt] = a;
  }
}
// max(x, 0) as one v_max_i32 on the bits: every float with the sign bit set
// (-0 included) is a negative integer and becomes +0, every other float is
// unchanged.  fmaxf(x, 0) adds a canonicalising v_max_f32 of the MFMA output.
__device__ __forceinline__ float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }
__device__ __forceinline__ v4s relu_bf16(v4f a) {
  const uint32_t p0 = bf2_pack(relu(a[0]), relu(a[1]));
  const uint32_t p1 = bf2_pack(relu(a[2]), relu(a[3]));
  return __builtin_bit_cast(v4s, make_uint2(p0, p1));
}
// LDS written by some lanes of a wave and read by others of the same wave:
// a wave's LDS instructions are performed in issue order, so keeping them in
// program order suffices.  Wavefront-scope fences wait on no counter; a
// workgroup fence would also wait for the wave's outstanding global loads
// and stores.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool DECODE>
__global__ void __launch_bounds__(kFwdBlock, kFwdBlocksPerCu) k_guide_fwd(GuideNet N, const float4* pos, const float4* dir,
                                                         uint32_t stride, const uint32_t* rows, uint32_t cap,
                                                         float* Z4, unsigned char* head) {
  __shared__ __attribute__((aligned(16))) uint16_t wts[kWLds];
  __shared__ __attribute__((aligned(16))) float bias[kBiasCount];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  const uint32_t M = min(*rows, cap);
  const uint32_t nTiles = (M + 15) / 16;
  if (blockIdx.x * kFwdWaves >= nTiles) return;  // block-uniform
  // The next tile's query is loaded while this one is computed: issued
  // before this tile's stores, waiting for it leaves them in flight (the
  // vector memory counter retires loads and stores in issue order).  Loads
  // and stores are unconditional so that the waits stay counts, not drains:
  // indexes past the last tile or row are clamped to it, a row past the
  // count computes the last row's query and stores the same bytes there.
  // The first query is loaded before the block barrier, which waits for it,
  // so the loop is entered with nothing in flight.
  const uint32_t tstep = gridDim.x * kFwdWaves;
  uint32_t tile = blockIdx.x * kFwdWaves + wave;
  auto query_row = [&](uint32_t t) { return (size_t)min(min(t, nTiles - 1) * 16 + (uint32_t)lr, M - 1) * stride; };
  float4 p4n = pos[query_row(tile)], d4n = dir[query_row(tile)];
  stage_weights_chain(N, wts, DECODE);
  stage_bias(N, bias, DECODE);
  __syncthreads();
  for (; tile < nTiles; tile += tstep) {
    // the weight fragments are read from LDS for every tile: without a fence
    // in the loop the compiler hoists all 34 of them (136 VGPRs) and spills
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t r0 = tile * 16;
    const float4 p4 = p4n, d4 = d4n;
    p4n = pos[query_row(tile + tstep)];
    d4n = dir[query_row(tile + tstep)];
    // the lane's chain (enc_chain_pair): a position coordinate in [0, 1]
    // through 12 octaves, or the direction's three coordinates through 4 each
    const float pc = lg == 0 ? p4.x : (lg == 1 ? p4.y : p4.z);
    const float lo = N.lo[lg < 3 ? lg : 0], hi = N.hi[lg < 3 ? lg : 0];
    const float px = fminf(1.0f, fmaxf(0.0f, (pc - lo) / (hi - lo)));
    float sn, cs;
    v4s in[6];
#pragma unroll
    for (int ks = 0; ks < 6; ks++) {  // staged columns 16 ks + 4 lg .. +3: chain elements 2 ks, 2 ks + 1
      uint32_t pr[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int m = 2 * ks + u;
        if (m == 0 || (lg == 3 && (m & 3) == 0)) {  // start of an axis: sincospi of the coordinate
          const float f = lg < 3 ? px : (m == 0 ? d4.x : (m == 4 ? d4.y : d4.z));
          sincospif(f, &sn, &cs);
        } else {  // the next octave: sin 2t = 2 sin t cos t, cos 2t = 1 - 2 sin^2 t
          const float s2 = 2.0f * sn * cs, c2 = fmaf(-2.0f * sn, sn, 1.0f);
          sn = s2;
          cs = c2;
        }
        pr[u] = bf2_pack(sn, cs);
      }
      in[ks] = __builtin_bit_cast(v4s, make_uint2(pr[0], pr[1]));
    }
    v4f acc[4];
    v4s h[4];
    fwd_layer<6, 4>(wts + kWOff[0], kGIn + kWPad, bias, in, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = relu_bf16(acc[k]);
    fwd_layer<4, 4>(wts + kWOff[1], kGHid + kWPad, bias + kGHid, h, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = relu_bf16(acc[k]);
    fwd_layer<4, 4>(wts + kWOff[2], kGHid + kWPad, bias + 2 * kGHid, h, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = relu_bf16(acc[k]);
    fwd_layer<4, 3>(wts + kWOff[3], kGHid + kWPad, bias + 3 * kGHid, h, acc, lr, lg);
    if (!DECODE) {  // row lr, outputs 16 mt + 4 lg .. +3
      const size_t orow = min(r0 + (uint32_t)lr, M - 1);
#pragma unroll
      for (int mt = 0; mt < 3; mt++)
        *reinterpret_cast<float4*>(Z4 + orow * kGOut + mt * 16 + 4 * lg) =
            make_float4(acc[mt][0], acc[mt][1], acc[mt][2], acc[mt][3]);
      continue;
    }
    {  // decode in the D fragments (decode_order): lane group lg holds lobes 2 lg and 2 lg + 1 of row lr
      unsigned char* hrow = head + (size_t)min(r0 + (uint32_t)lr, M - 1) * kGHead;
      const float l0 = acc[0][0], l1 = acc[1][1];
      float mx = fmaxf(l0, l1);
      mx = fmaxf(mx, __shfl_xor(mx, 16));
      mx = fmaxf(mx, __shfl_xor(mx, 32));
      const float e0 = __expf(l0 - mx), e1 = __expf(l1 - mx);
      // cumulative sums over the row's lane groups in lobe order (exclusive prefix of the pair sums)
      const float pair = e0 + e1;
      const float p1 = __shfl_xor(pair, 16), p2 = __shfl_xor(pair, 32), p3 = __shfl_xor(pair, 48);
      const float total = pair + p1 + p2 + p3;
      // groups lg = 0..3 hold pairs 0..3; the sum of the pairs before this group's
      const float before = (lg == 0) ? 0.0f : (lg == 1) ? p1 : (lg == 2) ? (p2 + p3) : (p1 + p2 + p3);
      // approximate reciprocals (1 ulp) in the decode: the sampler and the
      // density both read the stored values, so they stay consistent
      const float inv = __builtin_amdgcn_rcpf(total);
      const float cdf0 = (before + e0) * inv;
      const float cdf1 = lg == 3 ? 1.0f : (before + e0 + e1) * inv;
      const float cprev = before * inv;
      const float wts2[2] = {cdf0 - cprev, cdf1 - cdf0};
      const float cdfs[2] = {cdf0, cdf1};
      // lobe 2 lg: mean in slots 1-3, log kappa in slot 4; lobe 2 lg + 1: slots 6-8 and 9
      const float mu[2][3] = {{acc[0][1], acc[0][2], acc[0][3]}, {acc[1][2], acc[1][3], acc[2][0]}};
      const float lk[2] = {acc[1][0], acc[2][1]};
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int j = 2 * lg + u;
        const float m0 = mu[u][0], m1 = mu[u][1], m2 = mu[u][2];
        const float il = __builtin_amdgcn_rsqf(fmaxf(m0 * m0 + m1 * m1 + m2 * m2, 1e-12f));
        const float k = __expf(fminf(8.0f, fmaxf(-3.0f, lk[u])));
        reinterpret_cast<float4*>(hrow)[j] = make_float4(m0 * il, m1 * il, m2 * il, k);
        reinterpret_cast<float*>(hrow + kGHeadCw)[j] =
            wts2[u] * k * __builtin_amdgcn_rcpf(6.28318530717958647692f * (1.0f - __expf(-2.0f * k)));
        reinterpret_cast<float*>(hrow + kGHeadCdf)[j] = cdfs[u];
      }
      // the selection logit (slot 10 of group 0) to the row's four lanes, which store the same record
      const float al = __shfl(acc[2][2], lr);
      *reinterpret_cast<float4*>(hrow + kGHeadAlpha) =
          make_float4(kAlphaMin + kAlphaSpan / (1.0f + __expf(-al)), 0.0f, 0.0f, 0.0f);
    }
  }
}

hipError_t guide_infer(const GuideNet& N, const float4* pos, const float4* dir, uint32_t stride, const uint32_t* rows,
                       uint32_t cap, float* Z4, unsigned char* head, int n_cu, hipStream_t s) {
  // as many blocks as are resident at once, each grid-striding over the tiles
  const uint32_t tiles = (cap + 15) / 16;
  const int blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>((tiles + kFwdWaves - 1) / kFwdWaves,
                                                                   (uint32_t)n_cu * kFwdBlocksPerCu));
  if (head)
    hipLaunchKernelGGL(k_guide_fwd<true>, dim3(blocks), dim3(kFwdBlock), 0, s, N, pos, dir, stride, rows, cap, Z4,
                       head);
  else
    hipLaunchKernelGGL(k_guide_fwd<false>, dim3(blocks), dim3(kFwdBlock), 0, s, N, pos, dir, stride, rows, cap, Z4,
                       head);
  return hipGetLastError();
}

// ------------------------------------------------------------------ Adam
struct AdamArgs {
  float* p;  // fp32 master parameters (W1,W2,W3,W4 row-major in x out, then b1..b4)
  float* m;
  float* v;
  float* grad_out;  // optional: mean gradient
  const uint32_t* rows;
  uint32_t cap;
  float lr, b1, b2, eps, c1, c2;
  uint16_t* Wt[4];
  uint16_t* Wr[4];
};

__device__ __forceinline__ void param_loc(int idx, int& L, int& off, bool& isb) {
  const int wsz[4] = {kGIn * kGHid, kGHid * kGHid, kGHid * kGHid, kGHid * kGOut};
  int base = 0;
  for (L = 0; L < 4; L++) {
    if (idx < base + wsz[L]) {
      off = idx - base;
      isb = false;
      return;
    }
    base += wsz[L];
  }
  const int bsz[4] = {kGHid, kGHid, kGHid, kGOut};
  for (L = 0; L < 4; L++) {
    if (idx < base + bsz[L]) {
      off = idx - base;
      isb = true;
      return;
    }
    base += bsz[L];
  }
  L = -1;
}

// ------------------------------------------------------------------ fused training step
// One Adam step's gradients in one launch, one block of four waves per CU.
// Each wave takes 16 rows at a time: the encodings are built in registers,
// the forward and backward chains run in the transposed form of the
// inference kernel (D fragments of one layer are the B fragments of the
// next, activations and deltas stay in registers), and the loss and its
// gradient with respect to the 48 outputs are computed by four lanes per
// row.  The wave then stores its tile's inputs X^T, activations H_l^T and
// deltas D_l^T (bf16, feature-major) in LDS; after a block barrier the four
// waves split the 68 16x16 tiles of the weight gradients, each contracting
// over the block's 64 rows, and 240 threads sum the bias gradients.  Every
// block writes its own partial (kGParams floats) and loss sum; k_adam_reduce
// adds the partials in a fixed order, so the step is deterministic.
constexpr int kTrainWaves = 4;
constexpr int kTrainBlock = 64 * kTrainWaves;
// stored tile rows (features) of one wave: X 96, H1..H3, D1..D3 64 each, dZ 48
constexpr int kFX = 0, kFH1 = 96, kFH2 = 160, kFH3 = 224, kFD1 = 288, kFD2 = 352, kFD3 = 416, kFDZ = 480,
              kTileFeat = 528;
// backward weights in LDS: W_l [in][out] for l = 2, 3, 4 (rows padded by 8)
constexpr int kRPad = 8;
constexpr int kROff[3] = {0, kGHid * (kGHid + kRPad), 2 * kGHid * (kGHid + kRPad)};
constexpr int kRLds = kROff[2] + kGHid * (kGOut + kRPad);
// weight-gradient tiles: layer 1 4 x 6, layers 2, 3 4 x 4, layer 4 3 x 4 (out x in)
constexpr int kDwTiles = 24 + 16 + 16 + 12, kDwPerWave = kDwTiles / kTrainWaves;
static_assert(kDwTiles % kTrainWaves == 0, "weight-gradient tiles split evenly over the waves");
constexpr int kParamBase[4] = {0, kGIn * kGHid, kGIn * kGHid + kGHid * kGHid, kGIn * kGHid + 2 * kGHid * kGHid};
constexpr int kBiasBase = kGIn * kGHid + 2 * kGHid * kGHid + kGHid * kGOut;

struct TrainRows {
  const float4 *pos, *dir, *wdir;
  const float *pb, *wt;
};

__device__ __forceinline__ void dw_tile(int t, int& L, int& mo, int& mi) {
  if (t < 24) { L = 0; mo = t / 6; mi = t % 6; }
  else if (t < 40) { L = 1; mo = (t - 24) / 4; mi = (t - 24) % 4; }
  else if (t < 56) { L = 2; mo = (t - 40) / 4; mi = (t - 40) % 4; }
  else { L = 3; mo = (t - 56) / 4; mi = (t - 56) % 4; }
}

// A 16 x 16 D fragment (features 16 m + 4 lg + i, row lr) into a feature-major tile.
__device__ __forceinline__ void store_frag(uint16_t* tile, int f0, v4s v, int lr, int lg) {
#pragma unroll
  for (int i = 0; i < 4; i++) tile[(f0 + 4 * lg + i) * 16 + lr] = (uint16_t)v[i];
}

__global__ void __launch_bounds__(kTrainBlock) k_guide_train(GuideNet N, TrainRows T, const uint32_t* rows,
                                                             uint32_t cap, float* partial, float* loss_part) {
  __shared__ __attribute__((aligned(16))) uint16_t wts[kWLds];
  __shared__ __attribute__((aligned(16))) uint16_t wtr[kRLds];
  __shared__ __attribute__((aligned(16))) float bias[kBiasCount];
  __shared__ __attribute__((aligned(16))) uint16_t tiles[kTrainWaves][kTileFeat * 16];
  __shared__ __attribute__((aligned(16))) float zt[kTrainWaves][16 * kGOut];
  __shared__ float lsum[kTrainWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  const uint32_t M = min(*rows, cap);
  const uint32_t nIter = (M + 16 * kTrainWaves - 1) / (16 * kTrainWaves);
  stage_weights(N, wts);
  stage_bias(N, bias, false);
  for (int L = 1; L < 4; L++) {  // backward weights, [in][out]
    const int K = kLayerOut[L], ld = K + kRPad;
    for (int e = threadIdx.x; e < kGHid * K / 4; e += blockDim.x) {
      const int r = (4 * e) / K, col = (4 * e) % K;
      *reinterpret_cast<uint2*>(&wtr[kROff[L - 1] + r * ld + col]) =
          *reinterpret_cast<const uint2*>(N.Wr[L] + (size_t)r * K + col);
    }
  }
  v4f dw[kDwPerWave];
#pragma unroll
  for (int j = 0; j < kDwPerWave; j++) dw[j] = v4f{0, 0, 0, 0};
  float dbias = 0.0f, wloss = 0.0f;
  uint16_t* const tl = tiles[wave];
  float* const z = zt[wave];
  __syncthreads();
  for (uint32_t it = blockIdx.x; it < nIter; it += gridDim.x) {
    const uint32_t r0 = (it * kTrainWaves + wave) * 16;
    const uint32_t gr = min(r0 + (uint32_t)lr, M ? M - 1 : 0u);
    // ---- forward, transposed, in registers
    v4s in[6];
    {
      const float4 p4 = T.pos[gr], d4 = T.dir[gr];
      const Vec3 x = v3(p4.x, p4.y, p4.z), dd = v3(d4.x, d4.y, d4.z);
#pragma unroll
      for (int ks = 0; ks < 6; ks++) {
        const uint32_t p0 = guide_pair(N, x, dd, 8 * ks + 2 * lg), p1 = guide_pair(N, x, dd, 8 * ks + 2 * lg + 1);
        in[ks] = v4s{(short)(p0 & 0xffffu), (short)(p0 >> 16), (short)(p1 & 0xffffu), (short)(p1 >> 16)};
      }
    }
    v4f acc[4];
    v4s h1[4], h2[4], h3[4];
    fwd_layer<6, 4>(wts + kWOff[0], kGIn + kWPad, bias, in, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h1[k] = relu_bf16(acc[k]);
    fwd_layer<4, 4>(wts + kWOff[1], kGHid + kWPad, bias + kGHid, h1, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h2[k] = relu_bf16(acc[k]);
    fwd_layer<4, 4>(wts + kWOff[2], kGHid + kWPad, bias + 2 * kGHid, h2, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h3[k] = relu_bf16(acc[k]);
    fwd_layer<4, 3>(wts + kWOff[3], kGHid + kWPad, bias + 3 * kGHid, h3, acc, lr, lg);
#pragma unroll
    for (int mt = 0; mt < 3; mt++)
      *reinterpret_cast<float4*>(z + lr * kGOut + mt * 16 + 4 * lg) = make_float4(acc[mt][0], acc[mt][1], acc[mt][2], acc[mt][3]);
#pragma unroll
    for (int ks = 0; ks < 6; ks++) store_frag(tl, kFX + 16 * ks, in[ks], lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      store_frag(tl, kFH1 + 16 * k, h1[k], lr, lg);
      store_frag(tl, kFH2 + 16 * k, h2[k], lr, lg);
      store_frag(tl, kFH3 + 16 * k, h3[k], lr, lg);
    }
    wave_lds_sync();
    // ---- loss -wt log(alpha q + (1 - alpha) p_bsdf) and its gradient: four lanes per row, two lobes each
    {
      const int r = lane >> 2, sub = lane & 3;
      const uint32_t row = r0 + (uint32_t)r;
      const bool valid = row < M;
      const uint32_t rr = valid ? row : 0u;
      const float* zr = z + r * kGOut;
      const float4 w4 = T.wdir[rr];
      const float om[3] = {w4.x, w4.y, w4.z};
      const float pbr = T.pb[rr], wr = valid ? T.wt[rr] : 0.0f;
      float pi[2], v[2], mu[2][3], ml[2], kap[2], cj[2];
      bool clamped[2];
      float mx = fmaxf(zr[10 * sub], zr[10 * sub + 5]);
      mx = fmaxf(mx, __shfl_xor(mx, 1));
      mx = fmaxf(mx, __shfl_xor(mx, 2));
      float qs = 0.0f, es = 0.0f;
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int j = 2 * sub + u;
        pi[u] = expf(zr[5 * j] - mx);
        es += pi[u];
        const float m0 = zr[5 * j + 1], m1 = zr[5 * j + 2], m2 = zr[5 * j + 3];
        const float l = fmaxf(sqrtf(m0 * m0 + m1 * m1 + m2 * m2), 1e-6f);
        ml[u] = l;
        mu[u][0] = m0 / l;
        mu[u][1] = m1 / l;
        mu[u][2] = m2 / l;
        const float lk = zr[5 * j + 4];
        clamped[u] = lk < -3.0f || lk > 8.0f;
        kap[u] = expf(fminf(8.0f, fmaxf(-3.0f, lk)));
        cj[u] = mu[u][0] * om[0] + mu[u][1] * om[1] + mu[u][2] * om[2];
        v[u] = kap[u] / (6.28318530717958647692f * (1.0f - expf(-2.0f * kap[u]))) * expf(kap[u] * (cj[u] - 1.0f));
      }
      es += __shfl_xor(es, 1);
      es += __shfl_xor(es, 2);
#pragma unroll
      for (int u = 0; u < 2; u++) {
        pi[u] /= es;
        qs += pi[u] * v[u];
      }
      qs += __shfl_xor(qs, 1);
      qs += __shfl_xor(qs, 2);
      const float q = qs;
      const float sig = 1.0f / (1.0f + expf(-zr[40]));
      const float alpha = kAlphaMin + kAlphaSpan * sig;
      const float pm = fmaxf(alpha * q + (1.0f - alpha) * pbr, 1e-12f);
      const float gp = -wr / pm, dq = gp * alpha;
      uint16_t* dz = tl + kFDZ * 16;  // feature-major dZ^T
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int j = 2 * sub + u;
        const float dv = dq * pi[u], k = kap[u];
        const float e = expf(-2.0f * k);
        const float dlk = 1.0f / k - 2.0f * e / (1.0f - e) + (cj[u] - 1.0f);
        const float sc = dv * v[u] * k / ml[u];
        dz[(5 * j) * 16 + r] = f2bf(dq * pi[u] * (v[u] - q));
        for (int a = 0; a < 3; a++) dz[(5 * j + 1 + a) * 16 + r] = f2bf(sc * (om[a] - cj[u] * mu[u][a]));
        dz[(5 * j + 4) * 16 + r] = f2bf(clamped[u] ? 0.0f : dv * v[u] * dlk * k);
      }
      if (sub == 0) {
        dz[40 * 16 + r] = f2bf(gp * (q - pbr) * kAlphaSpan * sig * (1.0f - sig));
        wloss += valid ? -wr * logf(pm) : 0.0f;
      } else if (sub == 1) {
        for (int o = 41; o < kGOut; o++) dz[o * 16 + r] = 0;
      }
    }
    wave_lds_sync();
    // ---- backward deltas, transposed: D_l^T = (W_{l+1}^T D_{l+1}^T) * (H_l > 0)
    v4s dzb[3];
#pragma unroll
    for (int ks = 0; ks < 3; ks++)
#pragma unroll
      for (int s2 = 0; s2 < 4; s2++) dzb[ks][s2] = (short)tl[(kFDZ + 16 * ks + 4 * lg + s2) * 16 + lr];
    v4s dh[4];
    {
      const uint16_t* w4 = wtr + kROff[2];
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        // 48 outputs: two 32-wide steps, the second one's upper half zero
        // in both operands (every product of the network is the same
        // instruction, so each accumulator chain is one kind of MFMA)
        const uint16_t* wr = w4 + (mt * 16 + lr) * (kGOut + kRPad) + 4 * lg;
        const v4s z4 = v4s{0, 0, 0, 0};
        v4f a = mfma2(ld4(wr), ld4(wr + 16), dzb[0], dzb[1], v4f{0, 0, 0, 0});
        a = mfma2(ld4(wr + 32), z4, dzb[2], z4, a);
#pragma unroll
        for (int i = 0; i < 4; i++) dh[mt][i] = (short)f2bf(bf2f((uint16_t)h3[mt][i]) > 0.0f ? a[i] : 0.0f);
        store_frag(tl, kFD3 + 16 * mt, dh[mt], lr, lg);
      }
    }
#pragma unroll
    for (int step = 0; step < 2; step++) {  // D2 from D3 (W3), D1 from D2 (W2)
      const uint16_t* wl = wtr + kROff[1 - step];
      const v4s* hm = step == 0 ? h2 : h1;
      v4s nd[4];
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        v4f a = v4f{0, 0, 0, 0};
        const uint16_t* wr = wl + (mt * 16 + lr) * (kGHid + kRPad) + 4 * lg;
#pragma unroll
        for (int ks = 0; ks < 4; ks += 2) a = mfma2(ld4(wr + ks * 16), ld4(wr + ks * 16 + 16), dh[ks], dh[ks + 1], a);
#pragma unroll
        for (int i = 0; i < 4; i++) nd[mt][i] = (short)f2bf(bf2f((uint16_t)hm[mt][i]) > 0.0f ? a[i] : 0.0f);
        store_frag(tl, (step == 0 ? kFD2 : kFD1) + 16 * mt, nd[mt], lr, lg);
      }
#pragma unroll
      for (int mt = 0; mt < 4; mt++) dh[mt] = nd[mt];
    }
    __syncthreads();
    // ---- weight gradients: dW_l[out][in] += sum over the block's 64 rows of D_l^T[out][row] H_{l-1}^T[in][row]
#pragma unroll
    for (int j = 0; j < kDwPerWave; j++) {
      int L, mo, mi;
      dw_tile(wave * kDwPerWave + j, L, mo, mi);
      const int fD = L == 0 ? kFD1 : L == 1 ? kFD2 : L == 2 ? kFD3 : kFDZ;
      const int fH = L == 0 ? kFX : L == 1 ? kFH1 : L == 2 ? kFH2 : kFH3;
      v4f a = dw[j];
      static_assert(kTrainWaves % 2 == 0, "the block's rows in whole 32-row steps");
#pragma unroll
      for (int wv = 0; wv < kTrainWaves; wv += 2) {  // the rows of waves wv and wv + 1 as one 32-wide step
        const uint16_t *t0 = tiles[wv], *t1 = tiles[wv + 1];
        const int oD = (fD + 16 * mo + lr) * 16 + 4 * lg, oH = (fH + 16 * mi + lr) * 16 + 4 * lg;
        a = mfma2(ld4(t0 + oD), ld4(t1 + oD), ld4(t0 + oH), ld4(t1 + oH), a);
      }
      dw[j] = a;
    }
    if (threadIdx.x < 240) {  // bias gradients: feature threadIdx.x of D1 | D2 | D3 | dZ
      const int t = threadIdx.x;
      const int f = t < 64 ? kFD1 + t : t < 128 ? kFD2 + t - 64 : t < 192 ? kFD3 + t - 128 : kFDZ + t - 192;
      for (int wv = 0; wv < kTrainWaves; wv++)
        for (int r = 0; r < 16; r++) dbias += bf2f(tiles[wv][f * 16 + r]);
    }
    __syncthreads();
  }
  // ---- this block's partial: weights (D layout: out 16 mo + 4 lg + i, in 16 mi + lr), biases, loss
  float* out = partial + (size_t)blockIdx.x * kGParams;
#pragma unroll
  for (int j = 0; j < kDwPerWave; j++) {
    int L, mo, mi;
    dw_tile(wave * kDwPerWave + j, L, mo, mi);
    const int OUT = kLayerOut[L];
#pragma unroll
    for (int i = 0; i < 4; i++) out[kParamBase[L] + (16 * mi + lr) * OUT + 16 * mo + 4 * lg + i] = dw[j][i];
  }
  if (threadIdx.x < 240) out[kBiasBase + threadIdx.x] = dbias;
  for (int o = 1; o < 64; o <<= 1) wloss += __shfl_xor(wloss, o);
  if (lane == 0) lsum[wave] = wloss;
  __syncthreads();
  if (threadIdx.x == 0) loss_part[blockIdx.x] = lsum[0] + lsum[1] + lsum[2] + lsum[3];
}

// Mean gradient of a step from the training blocks' partials, then the Adam
// update of the FP32 masters and their BF16 copies.  A block takes 32
// parameters; its 8 lane groups of 32 add the partials b = g, g + 8, ... in
// ascending order, and the 8 group sums are added in group order: a fixed
// order, so the step is deterministic, with 8 loads in flight per parameter
// where one thread per parameter walked all partials.
constexpr int kAdamParams = 32, kAdamGroups = 8;
__global__ void __launch_bounds__(kAdamParams * kAdamGroups) k_adam_reduce(AdamArgs A, const float* partial,
                                                                           uint32_t nblocks) {
  __shared__ float red[kAdamGroups][kAdamParams];
  const int pi = threadIdx.x % kAdamParams, grp = threadIdx.x / kAdamParams;
  const int idx = blockIdx.x * kAdamParams + pi;
  float gs = 0.0f;
  if (idx < kGParams)
    for (uint32_t b = grp; b < nblocks; b += kAdamGroups) gs += partial[(size_t)b * kGParams + idx];
  red[grp][pi] = gs;
  __syncthreads();
  if (grp != 0 || idx >= kGParams) return;
  float g = red[0][pi];
#pragma unroll
  for (int k = 1; k < kAdamGroups; k++) g += red[k][pi];
  int L, off;
  bool isb;
  param_loc(idx, L, off, isb);
  const int IN = kLayerIn[L], OUT = kLayerOut[L];
  const uint32_t M = min(*A.rows, A.cap);
  g = M ? g / (float)M : 0.0f;
  if (A.grad_out) A.grad_out[idx] = g;
  float m = A.b1 * A.m[idx] + (1.0f - A.b1) * g;
  float v = A.b2 * A.v[idx] + (1.0f - A.b2) * g * g;
  A.m[idx] = m;
  A.v[idx] = v;
  float p = A.p[idx] - A.lr * (m / A.c1) / (sqrtf(v / A.c2) + A.eps);
  if (!isb && L == 3 && (off % OUT) >= 41) p = 0.0f;  // padding columns of the head stay zero
  if (isb && L == 3 && off >= 41) p = 0.0f;
  A.p[idx] = p;
  if (!isb) {
    const int i = off / OUT, j = off % OUT;
    const uint16_t h = f2bf(p);
    A.Wr[L][i * OUT + j] = h;
    A.Wt[L][j * IN + i] = h;
  }
}

// Master fp32 -> bf16 copies (after set_weights / init).
__global__ void k_guide_sync(AdamArgs A) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kGParams) return;
  int L, off;
  bool isb;
  param_loc(idx, L, off, isb);
  if (isb) return;
  const int IN = kLayerIn[L], OUT = kLayerOut[L];
  const int i = off / OUT, j = off % OUT;
  const uint16_t h = f2bf(A.p[idx]);
  A.Wr[L][i * OUT + j] = h;
  A.Wt[L][j * IN + i] = h;
}

// ------------------------------------------------------------------ state
static AdamArgs adam_args(GuideState& G) {
  AdamArgs A{};
  A.p = G.params.as<float>();
  A.m = G.adam_m.as<float>();
  A.v = G.adam_v.as<float>();
  for (int l = 0; l < 4; l++) {
    A.Wt[l] = G.Wt[l].as<uint16_t>();
    A.Wr[l] = G.Wr[l].as<uint16_t>();
  }
  return A;
}

int guide_reserve(rx_ctx* c, GuideState& G, uint32_t n) {
  hipError_t e = hipSuccess;
  auto E = [&](DBuf& b, size_t bytes) {
    if (e == hipSuccess) e = b.ensure(bytes);
  };
  E(G.tpart, (size_t)c->n_cu * kGParams * 4);  // one partial per training block (at most one per CU)
  E(G.loss_part, (size_t)c->n_cu * 4);
  if (n > G.cap) {
    const size_t N = n;
    E(G.Z4, N * std::max<size_t>(kGOut * 4, kGHead));  // raw outputs or decoded mixtures
    E(G.pos, N * 16);
    E(G.dir, N * 16);
    E(G.wdir, N * 16);
    E(G.pb, N * 4);
    E(G.wt, N * 4);
  }
  if (int r = hip_err(c, e, "guide buffers")) return r;
  G.cap = std::max(G.cap, n);
  return RX_OK;
}

int guide_reserve_store(rx_ctx* c, GuideState& G, uint32_t n) {
  if (n <= G.store_cap) return RX_OK;
  const size_t N = n;
  hipError_t e = hipSuccess;
  for (DBuf* b : {&G.spos, &G.sdir, &G.swdir}) e = e == hipSuccess ? b->ensure(N * 16) : e;
  for (DBuf* b : {&G.spb, &G.swt}) e = e == hipSuccess ? b->ensure(N * 4) : e;
  if (e == hipSuccess) e = G.srows.ensure(64);
  if (int r = hip_err(c, e, "guide sample store")) return r;
  G.store_cap = n;
  return RX_OK;
}

// One training step (k_guide_train + k_adam_reduce) on the rows in G.pos /
// G.dir / G.wdir / G.pb / G.wt with the row count at G.rows (device).  Loss
// partials stay on the device.
static int guide_train_device(rx_ctx* c, GuideState& G, uint32_t cap, float* grad_out) {
  hipStream_t s = c->stream;
  const uint32_t* rows = G.rows.as<uint32_t>();
  const uint32_t iters = (cap + 16 * kTrainWaves - 1) / (16 * kTrainWaves);
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(iters, (uint32_t)c->n_cu));
  const TrainRows T{G.pos.as<float4>(), G.dir.as<float4>(), G.wdir.as<float4>(), G.pb.as<float>(), G.wt.as<float>()};
  hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(kTrainBlock), 0, s, G.net(), T, rows, cap, G.tpart.as<float>(),
                     G.loss_part.as<float>());
  G.train_blocks = blocks;
  G.step++;
  AdamArgs A = adam_args(G);
  A.grad_out = grad_out;
  A.rows = rows;
  A.cap = cap;
  A.lr = G.prm.learning_rate;
  if ((G.prm.flags & RX_GUIDE_LR_DECAY) && G.decay_steps)  // geometric decay to a tenth over the warm-up
    A.lr *= powf(0.1f, std::min(1.0f, (float)(G.step - 1) / (float)G.decay_steps));
  A.b1 = G.prm.beta1;
  A.b2 = G.prm.beta2;
  A.eps = G.prm.eps;
  A.c1 = 1.0f - powf(A.b1, (float)G.step);
  A.c2 = 1.0f - powf(A.b2, (float)G.step);
  hipLaunchKernelGGL(k_adam_reduce, dim3((kGParams + kAdamParams - 1) / kAdamParams), dim3(kAdamParams * kAdamGroups),
                     0, s, A, G.tpart.as<float>(), blocks);
  return hip_err(c, hipGetLastError(), "guide train launch");
}

// Deterministic seeded init: uniform(-s, s) with s = sqrt(6 / fan_in) (He),
// the head scaled by 0.1 so the untrained mixture is broad and alpha ~ 0.5.
static void init_params(std::vector<float>& p, uint32_t seed) {
  p.assign(kGParams, 0.0f);
  uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
  auto rnd = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;
  };
  size_t base = 0;
  for (int l = 0; l < 4; l++) {
    const int IN = kLayerIn[l], OUT = kLayerOut[l];
    const double sc = std::sqrt(6.0 / IN) * (l == 3 ? 0.1 : 1.0);
    for (int i = 0; i < IN; i++)
      for (int j = 0; j < OUT; j++) p[base + (size_t)i * OUT + j] = (l == 3 && j >= 41) ? 0.0f : (float)((2 * rnd() - 1) * sc);
    base += (size_t)IN * OUT;
  }
}

// Train on rows the renderer already placed in G.pos/G.dir/G.wdir/G.pb/G.wt
// with the row count in G.rows (device); capacity `cap` was reserved.
int guide_train_rows(rx_ctx* c, uint32_t cap) { return guide_train_device(c, *c->guide, cap, nullptr); }

}  // namespace rx

using namespace rx;

extern "C" {

void rx_guide_state_free(rx_ctx* c) {
  if (c && c->guide) {
    delete c->guide;
    c->guide = nullptr;
  }
}

int rx_guide_init(rx_ctx* c, const rx_guide_params* prm) {
  if (!c || !prm) return RX_ERR_INVALID;
  if (prm->learning_rate <= 0 || prm->beta1 < 0 || prm->beta1 >= 1 || prm->beta2 < 0 || prm->beta2 >= 1)
    return set_err(c, RX_ERR_INVALID, "rx_guide_init: bad Adam parameters");
  (void)hipSetDevice(c->device);
  if (!c->guide) {
    c->guide = new (std::nothrow) GuideState;
    if (!c->guide) return RX_ERR_NO_MEMORY;
  }
  GuideState& G = *c->guide;
  G.prm = *prm;
  G.step = 0;
  hipStream_t s = c->stream;
  hipError_t e = hipSuccess;
  auto E = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  E(G.params.ensure(kGParams * 4));
  E(G.adam_m.ensure(kGParams * 4));
  E(G.adam_v.ensure(kGParams * 4));
  E(G.grad.ensure(kGParams * 4));
  E(G.rows.ensure(64));
  for (int l = 0; l < 4; l++) {
    E(G.Wt[l].ensure(kLayerIn[l] * kLayerOut[l] * 2));
    E(G.Wr[l].ensure(kLayerIn[l] * kLayerOut[l] * 2));
  }
  if (int r = hip_err(c, e, "guide alloc")) return r;
  std::vector<float> p;
  init_params(p, prm->seed);
  std::vector<float> zero(kGParams, 0.0f);
  if (int r = hip_err(c, G.params.upload(p.data(), kGParams * 4, s), "guide init")) return r;
  if (int r = hip_err(c, G.adam_m.upload(zero.data(), kGParams * 4, s), "guide init")) return r;
  if (int r = hip_err(c, G.adam_v.upload(zero.data(), kGParams * 4, s), "guide init")) return r;
  AdamArgs A = adam_args(G);
  hipLaunchKernelGGL(k_guide_sync, dim3((kGParams + 255) / 256), dim3(256), 0, s, A);
  if (int r = hip_err(c, hipGetLastError(), "guide sync")) return r;
  return hip_err(c, hipStreamSynchronize(s), "guide init sync");
}

int rx_guide_set_bounds(rx_ctx* c, const float* lo, const float* hi) {
  if (!c || !lo || !hi || !c->guide) return set_err(c, RX_ERR_STATE, "rx_guide_set_bounds: call rx_guide_init first");
  for (int a = 0; a < 3; a++)
    if (!(hi[a] > lo[a])) return set_err(c, RX_ERR_INVALID, "rx_guide_set_bounds: empty box");
  memcpy(c->guide->lo, lo, 12);
  memcpy(c->guide->hi, hi, 12);
  return RX_OK;
}

int rx_guide_set_weights(rx_ctx* c, const float* params) {
  if (!c || !params || !c->guide) return set_err(c, RX_ERR_STATE, "rx_guide_set_weights: call rx_guide_init first");
  GuideState& G = *c->guide;
  if (int r = hip_err(c, G.params.upload(params, kGParams * 4, c->stream), "set weights")) return r;
  AdamArgs A = adam_args(G);
  hipLaunchKernelGGL(k_guide_sync, dim3((kGParams + 255) / 256), dim3(256), 0, c->stream, A);
  return hip_err(c, hipStreamSynchronize(c->stream), "set weights sync");
}

int rx_guide_get_weights(rx_ctx* c, float* params) {
  if (!c || !params || !c->guide) return set_err(c, RX_ERR_STATE, "rx_guide_get_weights: call rx_guide_init first");
  if (int r = hip_err(c, hipMemcpyAsync(params, c->guide->params.p, kGParams * 4, hipMemcpyDeviceToHost, c->stream),
                      "get weights"))
    return r;
  return hip_err(c, hipStreamSynchronize(c->stream), "get weights sync");
}

static int upload_queries(rx_ctx* c, GuideState& G, uint32_t n, const float* pos, const float* dir) {
  std::vector<float> p4(4 * (size_t)n), d4(4 * (size_t)n);
  for (uint32_t i = 0; i < n; i++) {
    memcpy(&p4[4 * (size_t)i], pos + 3 * (size_t)i, 12);
    memcpy(&d4[4 * (size_t)i], dir + 3 * (size_t)i, 12);
  }
  hipStream_t s = c->stream;
  if (int r = hip_err(c, G.pos.upload(p4.data(), p4.size() * 4, s), "upload pos")) return r;
  if (int r = hip_err(c, G.dir.upload(d4.data(), d4.size() * 4, s), "upload dir")) return r;
  return hip_err(c, G.rows.upload(&n, 4, s), "upload rows");
}

int rx_guide_eval(rx_ctx* c, uint32_t n, const float* pos, const float* dir, float* out48) {
  if (!c || !c->guide) return set_err(c, RX_ERR_STATE, "rx_guide_eval: call rx_guide_init first");
  if (n == 0) return RX_OK;
  if (!pos || !dir || !out48) return set_err(c, RX_ERR_INVALID, "rx_guide_eval: null array");
  (void)hipSetDevice(c->device);
  GuideState& G = *c->guide;
  if (int r = guide_reserve(c, G, n)) return r;
  if (int r = upload_queries(c, G, n, pos, dir)) return r;
  if (int r = hip_err(c, guide_infer(G.net(), G.pos.as<float4>(), G.dir.as<float4>(), 1, G.rows.as<uint32_t>(), n,
                                     G.Z4.as<float>(), nullptr, c->n_cu, c->stream),
                      "guide inference"))
    return r;
  if (int r = hip_err(c, hipMemcpyAsync(out48, G.Z4.p, (size_t)n * kGOut * 4, hipMemcpyDeviceToHost, c->stream), "download"))
    return r;
  return hip_err(c, hipStreamSynchronize(c->stream), "eval sync");
}

int rx_guide_eval_mixture(rx_ctx* c, uint32_t n, const float* pos, const float* dir, float* out52) {
  if (!c || !c->guide) return set_err(c, RX_ERR_STATE, "rx_guide_eval_mixture: call rx_guide_init first");
  if (n == 0) return RX_OK;
  if (!pos || !dir || !out52) return set_err(c, RX_ERR_INVALID, "rx_guide_eval_mixture: null array");
  (void)hipSetDevice(c->device);
  GuideState& G = *c->guide;
  if (int r = guide_reserve(c, G, n)) return r;
  if (int r = upload_queries(c, G, n, pos, dir)) return r;
  if (int r = hip_err(c, guide_infer(G.net(), G.pos.as<float4>(), G.dir.as<float4>(), 1, G.rows.as<uint32_t>(), n,
                                     nullptr, G.Z4.as<unsigned char>(), c->n_cu, c->stream),
                      "guide inference"))
    return r;
  if (int r = hip_err(c, hipMemcpyAsync(out52, G.Z4.p, (size_t)n * kGHead, hipMemcpyDeviceToHost, c->stream), "download"))
    return r;
  return hip_err(c, hipStreamSynchronize(c->stream), "eval sync");
}

int rx_guide_train_batch(rx_ctx* c, uint32_t n, const float* pos, const float* dir, const float* omega,
                         const float* pdf_bsdf, const float* weight, float* loss, float* grad) {
  if (!c || !c->guide) return set_err(c, RX_ERR_STATE, "rx_guide_train_batch: call rx_guide_init first");
  if (n == 0) return RX_OK;
  if (!pos || !dir || !omega || !pdf_bsdf || !weight) return set_err(c, RX_ERR_INVALID, "rx_guide_train_batch: null array");
  (void)hipSetDevice(c->device);
  GuideState& G = *c->guide;
  if (int r = guide_reserve(c, G, n)) return r;
  if (int r = upload_queries(c, G, n, pos, dir)) return r;
  std::vector<float> w4(4 * (size_t)n);
  for (uint32_t i = 0; i < n; i++) memcpy(&w4[4 * (size_t)i], omega + 3 * (size_t)i, 12);
  hipStream_t s = c->stream;
  if (int r = hip_err(c, G.wdir.upload(w4.data(), w4.size() * 4, s), "upload omega")) return r;
  if (int r = hip_err(c, G.pb.upload(pdf_bsdf, 4 * (size_t)n, s), "upload pb")) return r;
  if (int r = hip_err(c, G.wt.upload(weight, 4 * (size_t)n, s), "upload weight")) return r;
  if (int r = guide_train_device(c, G, n, grad ? G.grad.as<float>() : nullptr)) return r;
  if (loss) {
    std::vector<float> lp(G.train_blocks);
    if (int r = hip_err(c, hipMemcpyAsync(lp.data(), G.loss_part.p, lp.size() * 4, hipMemcpyDeviceToHost, s), "loss"))
      return r;
    if (int r = hip_err(c, hipStreamSynchronize(s), "loss sync")) return r;
    double t = 0;
    for (float v : lp) t += v;
    *loss = (float)(t / n);
  }
  if (grad) {
    if (int r = hip_err(c, hipMemcpyAsync(grad, G.grad.p, kGParams * 4, hipMemcpyDeviceToHost, s), "grad")) return r;
  }
  return hip_err(c, hipStreamSynchronize(s), "train sync");
}

}  // extern "C"
