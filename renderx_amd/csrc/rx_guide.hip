// Path-guiding network on gfx950 matrix cores: BF16 MFMA forward, backward
// and weight-gradient kernels with FP32 accumulation, deterministic gradient
// reduction and an FP32 Adam step.  No counterpart in the reference.
//
// Every product in the network has one dimension of 16 rows per wavefront
// (a tile of queries), so all of it maps onto v_mfma_f32_16x16x16_bf16:
//   A fragment: lane l holds A[l%16][4(l/16) .. 4(l/16)+3]
//   B fragment: lane l holds B[4(l/16) .. 4(l/16)+3][l%16]
//   D fragment: lane l holds D[4(l/16)+i][l%16], i = 0..3
// Activations move between layers through a per-wave LDS tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "rx_context.h"
#include "rx_guide.h"

namespace rx {

constexpr int kDwChunk = 256;  // rows per weight-gradient partial

// ------------------------------------------------------------------ forward
// X [M x 96] bf16 -> A1..A3 [M x 64] bf16 (post-ReLU), Z4 [M x 48] fp32.
// Each block stages the four weight matrices in LDS once (rows padded by 8
// bf16 so that a fragment load's 16 rows fall in different banks), then its
// waves run 16-row tiles: the weights are read from HBM/L2 once per block
// instead of once per tile.
__global__ void __launch_bounds__(256) k_mlp_fwd(GuideNet N, const uint16_t* X, const uint32_t* rows, uint32_t cap,
                                                 uint16_t* A1, uint16_t* A2, uint16_t* A3, float* Z4) {
  __shared__ __attribute__((aligned(16))) uint16_t act[2][4][16 * 64];
  __shared__ __attribute__((aligned(16))) uint16_t wts[kWLds];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  const uint32_t M = min(*rows, cap);
  const uint32_t nTiles = (M + 15) / 16;
  if (blockIdx.x * 4 >= nTiles) return;  // no tile for this block (uniform)
  stage_weights(N, wts);
  __syncthreads();
  uint16_t* const outs[3] = {A1, A2, A3};
  for (uint32_t tb = blockIdx.x * 4; tb < nTiles; tb += gridDim.x * 4) {
    const uint32_t tile = tb + wave;
    const uint32_t row0 = tile * 16;
    const uint32_t rowA = min(row0 + (uint32_t)lr, M ? M - 1 : 0u);
    const bool live = tile < nTiles;
    for (int L = 0; L < 4; L++) {
      const int K = kLayerIn[L], NO = kLayerOut[L], NT = NO / 16;
      v4f acc[4];
      for (int nt = 0; nt < 4; nt++) acc[nt] = v4f{0, 0, 0, 0};
      const uint16_t* ain = (L == 0) ? nullptr : act[(L - 1) & 1][wave];
      const uint16_t* wl = wts + kWOff[L];
      for (int ks = 0; ks < K / 16; ks++) {
        v4s a = live ? ((L == 0) ? ld4(X + (size_t)rowA * kGIn + ks * 16 + 4 * lg) : ld4(ain + lr * 64 + ks * 16 + 4 * lg))
                     : v4s{0, 0, 0, 0};
        for (int nt = 0; nt < NT; nt++) {
          v4s b = ld4(wl + (nt * 16 + lr) * (K + kWPad) + ks * 16 + 4 * lg);
          acc[nt] = mfma(a, b, acc[nt]);
        }
      }
      for (int nt = 0; nt < NT; nt++) {
        const int col = nt * 16 + lr;
        const float bias = N.b[L][col];
        for (int i = 0; i < 4; i++) {
          const int row = 4 * lg + i;
          const float z = acc[nt][i] + bias;
          const uint32_t gr = row0 + row;
          if (L < 3) {
            const uint16_t h = f2bf(fmaxf(z, 0.0f));
            act[L & 1][wave][row * 64 + col] = h;
            if (live && gr < M && outs[L]) outs[L][(size_t)gr * kGHid + col] = h;
          } else if (live && gr < M) {
            Z4[(size_t)gr * kGOut + col] = z;
          }
        }
      }
      __syncthreads();
    }
  }
}

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
// renderer's form, GuideHead), staged through LDS so that rows are written
// whole.  Query row r is pos[r * stride], dir[r * stride].
constexpr int kFwdBlock = 512;  // 8 waves share one copy of the weights
constexpr int kFwdWaves = kFwdBlock / 64;

// One layer: acc[mt] = W[16 mt .. , :] * in + bias for the layer's MT output tiles.
template <int KS, int MT>
__device__ __forceinline__ void fwd_layer(const uint16_t* w, int ldw, const float* bias, const v4s* in, v4f* acc,
                                          int lr, int lg) {
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
    v4f a = v4f{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) a = mfma(ld4(w + (mt * 16 + lr) * ldw + ks * 16 + 4 * lg), in[ks], a);
#pragma unroll
    for (int i = 0; i < 4; i++) a[i] += bias[mt * 16 + 4 * lg + i];
    acc[mt] = a;
  }
}
__device__ __forceinline__ v4s relu_bf16(v4f a) {
  v4s r;
#pragma unroll
  for (int i = 0; i < 4; i++) r[i] = (short)f2bf(fmaxf(a[i], 0.0f));
  return r;
}
// LDS written by some lanes of a wave, read by others: the wave's LDS
// operations complete in order once issued.
__device__ __forceinline__ void wave_lds_sync() {
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
}

template <bool DECODE>
__global__ void __launch_bounds__(kFwdBlock) k_guide_fwd(GuideNet N, const float4* pos, const float4* dir,
                                                         uint32_t stride, const uint32_t* rows, uint32_t cap,
                                                         float* Z4, unsigned char* head) {
  __shared__ __attribute__((aligned(16))) uint16_t wts[kWLds];
  __shared__ __attribute__((aligned(16))) float zt[DECODE ? kFwdWaves : 1][16 * kGOut];    // a tile's outputs
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  const uint32_t M = min(*rows, cap);
  const uint32_t nTiles = (M + 15) / 16;
  if (blockIdx.x * kFwdWaves >= nTiles) return;  // block-uniform
  stage_weights(N, wts);
  __syncthreads();
  for (uint32_t tile = blockIdx.x * kFwdWaves + wave; tile < nTiles; tile += gridDim.x * kFwdWaves) {
    const uint32_t r0 = tile * 16;
    const uint32_t gr = min(r0 + (uint32_t)lr, M - 1);
    const float4 p4 = pos[(size_t)gr * stride], d4 = dir[(size_t)gr * stride];
    const Vec3 x = v3(p4.x, p4.y, p4.z), dd = v3(d4.x, d4.y, d4.z);
    v4s in[6];
#pragma unroll
    for (int ks = 0; ks < 6; ks++) {  // features 16 ks + 4 lg .. +3 = pairs 8 ks + 2 lg, +1
      const uint32_t p0 = guide_pair(N, x, dd, 8 * ks + 2 * lg), p1 = guide_pair(N, x, dd, 8 * ks + 2 * lg + 1);
      in[ks] = v4s{(short)(p0 & 0xffffu), (short)(p0 >> 16), (short)(p1 & 0xffffu), (short)(p1 >> 16)};
    }
    v4f acc[4];
    v4s h[4];
    fwd_layer<6, 4>(wts + kWOff[0], kGIn + kWPad, N.b[0], in, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = relu_bf16(acc[k]);
    fwd_layer<4, 4>(wts + kWOff[1], kGHid + kWPad, N.b[1], h, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = relu_bf16(acc[k]);
    fwd_layer<4, 4>(wts + kWOff[2], kGHid + kWPad, N.b[2], h, acc, lr, lg);
#pragma unroll
    for (int k = 0; k < 4; k++) h[k] = relu_bf16(acc[k]);
    fwd_layer<4, 3>(wts + kWOff[3], kGHid + kWPad, N.b[3], h, acc, lr, lg);
    if (!DECODE) {  // row lr, outputs 16 mt + 4 lg .. +3
      if (r0 + (uint32_t)lr < M)
#pragma unroll
        for (int mt = 0; mt < 3; mt++)
          *reinterpret_cast<float4*>(Z4 + (size_t)(r0 + lr) * kGOut + mt * 16 + 4 * lg) =
              make_float4(acc[mt][0], acc[mt][1], acc[mt][2], acc[mt][3]);
      continue;
    }
    float* z = zt[DECODE ? wave : 0];
#pragma unroll
    for (int mt = 0; mt < 3; mt++)
      *reinterpret_cast<float4*>(z + lr * kGOut + mt * 16 + 4 * lg) =
          make_float4(acc[mt][0], acc[mt][1], acc[mt][2], acc[mt][3]);
    wave_lds_sync();
    {  // four lanes per row, two lobes each, written to the row's lobe records directly
      const int r = lane >> 2, sub = lane & 3;
      const float* zr = z + r * kGOut;
      const bool out = r0 + (uint32_t)r < M;
      unsigned char* hrow = head + (size_t)(r0 + r) * kGHead;
      GuideLobe* lob = reinterpret_cast<GuideLobe*>(hrow);
      const int j0 = 2 * sub, j1 = 2 * sub + 1;
      const float l0 = zr[5 * j0], l1 = zr[5 * j1];
      float mx = fmaxf(l0, l1);
      mx = fmaxf(mx, __shfl_xor(mx, 1));
      mx = fmaxf(mx, __shfl_xor(mx, 2));
      const float e0 = __expf(l0 - mx), e1 = __expf(l1 - mx);
      float sum = e0 + e1;
      sum += __shfl_xor(sum, 1);
      sum += __shfl_xor(sum, 2);
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int j = u ? j1 : j0;
        const float m0 = zr[5 * j + 1], m1 = zr[5 * j + 2], m2 = zr[5 * j + 3];
        const float l = fmaxf(sqrtf(m0 * m0 + m1 * m1 + m2 * m2), 1e-6f);
        GuideLobe L;
        L.mu[0] = (half_t)(m0 / l);
        L.mu[1] = (half_t)(m1 / l);
        L.mu[2] = (half_t)(m2 / l);
        L.kappa = (half_t)__expf(fminf(8.0f, fmaxf(-3.0f, zr[5 * j + 4])));
        const float k = (float)L.kappa;  // the normaliser of the stored concentration
        L.pi = (u ? e1 : e0) / sum;
        L.norm = k / (6.28318530717958647692f * (1.0f - __expf(-2.0f * k)));
        if (out) lob[j] = L;
      }
      if (sub == 0 && out)
        *reinterpret_cast<float4*>(hrow + 8 * sizeof(GuideLobe)) =
            make_float4(kAlphaMin + kAlphaSpan / (1.0f + __expf(-zr[40])), 0.0f, 0.0f, 0.0f);
    }
    wave_lds_sync();  // the tile's outputs are read before the next tile overwrites them
  }
}

hipError_t guide_infer(const GuideNet& N, const float4* pos, const float4* dir, uint32_t stride, const uint32_t* rows,
                       uint32_t cap, float* Z4, unsigned char* head, int n_cu, hipStream_t s) {
  const uint32_t tiles = (cap + 15) / 16;
  const int blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>((tiles + kFwdWaves - 1) / kFwdWaves,
                                                                   (uint32_t)n_cu * 2));
  if (head)
    hipLaunchKernelGGL(k_guide_fwd<true>, dim3(blocks), dim3(kFwdBlock), 0, s, N, pos, dir, stride, rows, cap, Z4,
                       head);
  else
    hipLaunchKernelGGL(k_guide_fwd<false>, dim3(blocks), dim3(kFwdBlock), 0, s, N, pos, dir, stride, rows, cap, Z4,
                       head);
  return hipGetLastError();
}

// ------------------------------------------------------------------ backward deltas
// dZ4 [M x 48] bf16 -> D3, D2, D1 [M x 64] bf16:  D_l = (D_{l+1} W_{l+1}^T) * (A_l > 0)
__global__ void __launch_bounds__(256) k_mlp_bwd(GuideNet N, const uint16_t* dZ4, const uint32_t* rows, uint32_t cap,
                                                 const uint16_t* A1, const uint16_t* A2, const uint16_t* A3,
                                                 uint16_t* D1, uint16_t* D2, uint16_t* D3) {
  __shared__ __attribute__((aligned(16))) uint16_t dl[2][4][16 * 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  const uint32_t M = min(*rows, cap);
  const uint32_t nTiles = (M + 15) / 16;
  const uint16_t* acts[3] = {A1, A2, A3};
  uint16_t* outs[3] = {D1, D2, D3};
  for (uint32_t tb = blockIdx.x * 4; tb < nTiles; tb += gridDim.x * 4) {
    const uint32_t tile = tb + wave;
    const uint32_t row0 = tile * 16;
    const uint32_t rowA = min(row0 + (uint32_t)lr, M ? M - 1 : 0u);
    const bool live = tile < nTiles;
    // step s: from layer L+1's delta (K = out of layer L+1) to layer L's delta
    for (int s = 0; s < 3; s++) {
      const int L = 2 - s;             // producing D_{L+1} in 1-based terms: D3, D2, D1
      const int K = kLayerOut[L + 1];  // 48 for the first step, then 64
      v4f acc[4];
      for (int nt = 0; nt < 4; nt++) acc[nt] = v4f{0, 0, 0, 0};
      const uint16_t* din = dl[(s + 1) & 1][wave];
      for (int ks = 0; ks < K / 16; ks++) {
        v4s a = live ? ((s == 0) ? ld4(dZ4 + (size_t)rowA * kGOut + ks * 16 + 4 * lg) : ld4(din + lr * 64 + ks * 16 + 4 * lg))
                     : v4s{0, 0, 0, 0};
        for (int nt = 0; nt < 4; nt++) {
          v4s b = ld4(N.Wr[L + 1] + (size_t)(nt * 16 + lr) * K + ks * 16 + 4 * lg);
          acc[nt] = mfma(a, b, acc[nt]);
        }
      }
      for (int nt = 0; nt < 4; nt++) {
        const int col = nt * 16 + lr;
        for (int i = 0; i < 4; i++) {
          const int row = 4 * lg + i;
          const uint32_t gr = row0 + row;
          float v = 0.0f;
          if (live && gr < M) {
            const float a = bf2f(acts[L][(size_t)gr * kGHid + col]);
            v = a > 0.0f ? acc[nt][i] : 0.0f;
          }
          const uint16_t h = f2bf(v);
          dl[s & 1][wave][row * 64 + col] = h;
          if (live && gr < M) outs[L][(size_t)gr * kGHid + col] = h;
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ weight gradients
// partial[c][m*OUT + n] = sum over chunk c's rows of Ain[r][m] * D[r][n];
// partial[c][IN*OUT + n] = sum of D[r][n].  One wave per chunk of kDwChunk rows.
template <int IN, int OUT>
__global__ void __launch_bounds__(256) k_mlp_dw(const uint16_t* Ain, const uint16_t* D, const uint32_t* rows,
                                                uint32_t cap, float* partial) {
  constexpr int MT = IN / 16, NT = OUT / 16, P = IN * OUT + OUT;
  __shared__ __attribute__((aligned(16))) uint16_t ta[4][16 * IN];
  __shared__ __attribute__((aligned(16))) uint16_t td[4][16 * OUT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lg = lane >> 4;
  const uint32_t M = min(*rows, cap);
  const uint32_t nChunks = (M + kDwChunk - 1) / kDwChunk;
  for (uint32_t cb = blockIdx.x * 4; cb < nChunks; cb += gridDim.x * 4) {
    const uint32_t chunk = cb + wave;
    v4f acc[MT][NT];
    for (int m = 0; m < MT; m++)
      for (int n = 0; n < NT; n++) acc[m][n] = v4f{0, 0, 0, 0};
    float dbs = 0.0f;
    for (int kt = 0; kt < kDwChunk / 16; kt++) {
      const uint32_t r0 = chunk * kDwChunk + kt * 16;
      // stage 16 rows of Ain and D (zero beyond M)
      for (int e = lane; e < 16 * IN / 4; e += 64) {
        const int r = (e * 4) / IN, c = (e * 4) % IN;
        uint2 v = make_uint2(0, 0);
        if (chunk < nChunks && r0 + r < M) v = *reinterpret_cast<const uint2*>(Ain + (size_t)(r0 + r) * IN + c);
        *reinterpret_cast<uint2*>(&ta[wave][r * IN + c]) = v;
      }
      for (int e = lane; e < 16 * OUT / 4; e += 64) {
        const int r = (e * 4) / OUT, c = (e * 4) % OUT;
        uint2 v = make_uint2(0, 0);
        if (chunk < nChunks && r0 + r < M) v = *reinterpret_cast<const uint2*>(D + (size_t)(r0 + r) * OUT + c);
        *reinterpret_cast<uint2*>(&td[wave][r * OUT + c]) = v;
      }
      __syncthreads();
      v4s b[NT];
      for (int n = 0; n < NT; n++)
        for (int s = 0; s < 4; s++) b[n][s] = (short)td[wave][(4 * lg + s) * OUT + n * 16 + lr];
      for (int m = 0; m < MT; m++) {
        v4s a;
        for (int s = 0; s < 4; s++) a[s] = (short)ta[wave][(4 * lg + s) * IN + m * 16 + lr];
        for (int n = 0; n < NT; n++) acc[m][n] = mfma(a, b[n], acc[m][n]);
      }
      if (lane < OUT)
        for (int r = 0; r < 16; r++) dbs += bf2f(td[wave][r * OUT + lane]);
      __syncthreads();
    }
    if (chunk < nChunks) {
      float* out = partial + (size_t)chunk * P;
      for (int m = 0; m < MT; m++)
        for (int n = 0; n < NT; n++)
          for (int i = 0; i < 4; i++) out[(m * 16 + 4 * lg + i) * OUT + n * 16 + lr] = acc[m][n][i];
      if (lane < OUT) out[IN * OUT + lane] = dbs;
    }
  }
}

// ------------------------------------------------------------------ loss
// Per row: decode the head, the one-sample-MIS mixture density
// p = alpha q(w) + (1 - alpha) p_bsdf(w), loss -wt log p and its gradient with
// respect to the 48 outputs (bf16).  loss_part[block] = block sum (fixed order).
__global__ void __launch_bounds__(256) k_guide_loss(const float* Z4, const float4* wdir, const float* pb,
                                                    const float* wt, const uint32_t* rows, uint32_t cap,
                                                    uint16_t* dZ4, float* loss_part) {
  __shared__ float red[256];
  const uint32_t M = min(*rows, cap);
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  float loss = 0.0f;
  if (r < M) {
    const float* z = Z4 + (size_t)r * kGOut;
    float zz[kGOut];
    for (int c = 0; c < kGOut; c++) zz[c] = z[c];
    float mx = zz[0];
    for (int j = 1; j < kGLobes; j++) mx = fmaxf(mx, zz[5 * j]);
    float pi[kGLobes], s = 0.0f;
    for (int j = 0; j < kGLobes; j++) {
      pi[j] = expf(zz[5 * j] - mx);
      s += pi[j];
    }
    const float4 w4 = wdir[r];
    const float om[3] = {w4.x, w4.y, w4.z};
    float v[kGLobes], mu[kGLobes][3], ml[kGLobes], kap[kGLobes], cj[kGLobes];
    bool clamped[kGLobes];
    float q = 0.0f;
    for (int j = 0; j < kGLobes; j++) {
      pi[j] /= s;
      float m0 = zz[5 * j + 1], m1 = zz[5 * j + 2], m2 = zz[5 * j + 3];
      float l = fmaxf(sqrtf(m0 * m0 + m1 * m1 + m2 * m2), 1e-6f);
      ml[j] = l;
      mu[j][0] = m0 / l;
      mu[j][1] = m1 / l;
      mu[j][2] = m2 / l;
      float lk = zz[5 * j + 4];
      clamped[j] = lk < -3.0f || lk > 8.0f;
      kap[j] = expf(fminf(8.0f, fmaxf(-3.0f, lk)));
      cj[j] = mu[j][0] * om[0] + mu[j][1] * om[1] + mu[j][2] * om[2];
      v[j] = kap[j] / (6.28318530717958647692f * (1.0f - expf(-2.0f * kap[j]))) * expf(kap[j] * (cj[j] - 1.0f));
      q += pi[j] * v[j];
    }
    const float sig = 1.0f / (1.0f + expf(-zz[40]));
    const float alpha = kAlphaMin + kAlphaSpan * sig;
    const float pbr = pb[r], wr = wt[r];
    const float pm = fmaxf(alpha * q + (1.0f - alpha) * pbr, 1e-12f);
    loss = -wr * logf(pm);
    float g[kGOut];
    for (int c = 0; c < kGOut; c++) g[c] = 0.0f;
    const float gp = -wr / pm, dq = gp * alpha;
    g[40] = gp * (q - pbr) * kAlphaSpan * sig * (1.0f - sig);
    for (int j = 0; j < kGLobes; j++) {
      g[5 * j] = dq * pi[j] * (v[j] - q);
      const float dv = dq * pi[j], k = kap[j];
      const float e = expf(-2.0f * k);
      const float dlk = 1.0f / k - 2.0f * e / (1.0f - e) + (cj[j] - 1.0f);
      g[5 * j + 4] = clamped[j] ? 0.0f : dv * v[j] * dlk * k;
      const float sc = dv * v[j] * k / ml[j];
      for (int a = 0; a < 3; a++) g[5 * j + 1 + a] = sc * (om[a] - cj[j] * mu[j][a]);
    }
    for (int c = 0; c < kGOut; c++) dZ4[(size_t)r * kGOut + c] = f2bf(g[c]);
  }
  red[threadIdx.x] = loss;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------ Adam
struct AdamArgs {
  float* p;  // fp32 master parameters (W1,W2,W3,W4 row-major in x out, then b1..b4)
  float* m;
  float* v;
  float* grad_out;  // optional: mean gradient
  const float* partial[4];
  uint32_t nchunks;
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

__global__ void k_adam(AdamArgs A) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kGParams) return;
  int L, off;
  bool isb;
  param_loc(idx, L, off, isb);
  const int IN = kLayerIn[L], OUT = kLayerOut[L], P = IN * OUT + OUT;
  const int poff = isb ? IN * OUT + off : off;
  const uint32_t M = min(*A.rows, A.cap);
  const uint32_t nch = (M + kDwChunk - 1) / kDwChunk;
  float g = 0.0f;
  for (uint32_t c = 0; c < nch; c++) g += A.partial[L][(size_t)c * P + poff];
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

// Encode queries given as positions and directions (test and training path).
__global__ void k_guide_encode(GuideNet N, const float4* pos, const float4* dir, const uint32_t* rows, uint32_t cap,
                               uint16_t* X) {
  const uint32_t M = min(*rows, cap);
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < M; r += gridDim.x * blockDim.x) {
    float4 p = pos[r], d = dir[r];
    guide_encode(N, v3(p.x, p.y, p.z), v3(d.x, d.y, d.z), X + (size_t)r * kGIn);
  }
}

hipError_t guide_forward(const GuideNet& N, const uint16_t* X, const uint32_t* rows, uint32_t cap, uint16_t* A1,
                         uint16_t* A2, uint16_t* A3, float* Z4, int n_cu, hipStream_t s) {
  const uint32_t tiles = (cap + 15) / 16;
  const int blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>((tiles + 3) / 4, (uint32_t)n_cu * 8));
  hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks), dim3(256), 0, s, N, X, rows, cap, A1, A2, A3, Z4);
  return hipGetLastError();
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
    A.partial[l] = G.partial[l].as<float>();
  }
  return A;
}

int guide_reserve(rx_ctx* c, GuideState& G, uint32_t n) {
  if (n <= G.cap) return RX_OK;
  const size_t N = n;
  hipError_t e = hipSuccess;
  auto E = [&](DBuf& b, size_t bytes) {
    if (e == hipSuccess) e = b.ensure(bytes);
  };
  E(G.X, N * kGIn * 2);
  E(G.A1, N * kGHid * 2);
  E(G.A2, N * kGHid * 2);
  E(G.A3, N * kGHid * 2);
  E(G.Z4, N * kGOut * 4);
  E(G.dZ4, N * kGOut * 2);
  E(G.D1, N * kGHid * 2);
  E(G.D2, N * kGHid * 2);
  E(G.D3, N * kGHid * 2);
  const size_t nch = (N + kDwChunk - 1) / kDwChunk;
  for (int l = 0; l < 4; l++) E(G.partial[l], nch * (kLayerIn[l] * kLayerOut[l] + kLayerOut[l]) * 4);
  E(G.loss_part, ((N + 255) / 256) * 4);
  E(G.pos, N * 16);
  E(G.dir, N * 16);
  E(G.wdir, N * 16);
  E(G.pb, N * 4);
  E(G.wt, N * 4);
  if (int r = hip_err(c, e, "guide buffers")) return r;
  G.cap = n;
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

// One training step on rows already in G.X / G.wdir / G.pb / G.wt with the
// row count at G.rows (device).  Loss partials stay on the device.
static int guide_train_device(rx_ctx* c, GuideState& G, uint32_t cap, float* grad_out) {
  hipStream_t s = c->stream;
  GuideNet N = G.net();
  const uint32_t* rows = G.rows.as<uint32_t>();
  if (int r = hip_err(c, guide_forward(N, G.X.as<uint16_t>(), rows, cap, G.A1.as<uint16_t>(), G.A2.as<uint16_t>(),
                                       G.A3.as<uint16_t>(), G.Z4.as<float>(), c->n_cu, s),
                      "guide forward"))
    return r;
  hipLaunchKernelGGL(k_guide_loss, dim3((cap + 255) / 256), dim3(256), 0, s, G.Z4.as<float>(), G.wdir.as<float4>(),
                     G.pb.as<float>(), G.wt.as<float>(), rows, cap, G.dZ4.as<uint16_t>(), G.loss_part.as<float>());
  const uint32_t tiles = (cap + 15) / 16;
  const int bb = (int)std::max<uint32_t>(1, std::min<uint32_t>((tiles + 3) / 4, (uint32_t)c->n_cu * 8));
  hipLaunchKernelGGL(k_mlp_bwd, dim3(bb), dim3(256), 0, s, N, G.dZ4.as<uint16_t>(), rows, cap, G.A1.as<uint16_t>(),
                     G.A2.as<uint16_t>(), G.A3.as<uint16_t>(), G.D1.as<uint16_t>(), G.D2.as<uint16_t>(),
                     G.D3.as<uint16_t>());
  const uint32_t nch = (cap + kDwChunk - 1) / kDwChunk;
  const int db = (int)std::max<uint32_t>(1, (nch + 3) / 4);
  hipLaunchKernelGGL((k_mlp_dw<kGIn, kGHid>), dim3(db), dim3(256), 0, s, G.X.as<uint16_t>(), G.D1.as<uint16_t>(),
                     rows, cap, G.partial[0].as<float>());
  hipLaunchKernelGGL((k_mlp_dw<kGHid, kGHid>), dim3(db), dim3(256), 0, s, G.A1.as<uint16_t>(), G.D2.as<uint16_t>(),
                     rows, cap, G.partial[1].as<float>());
  hipLaunchKernelGGL((k_mlp_dw<kGHid, kGHid>), dim3(db), dim3(256), 0, s, G.A2.as<uint16_t>(), G.D3.as<uint16_t>(),
                     rows, cap, G.partial[2].as<float>());
  hipLaunchKernelGGL((k_mlp_dw<kGHid, kGOut>), dim3(db), dim3(256), 0, s, G.A3.as<uint16_t>(), G.dZ4.as<uint16_t>(),
                     rows, cap, G.partial[3].as<float>());
  G.step++;
  AdamArgs A = adam_args(G);
  A.grad_out = grad_out;
  A.nchunks = nch;
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
  hipLaunchKernelGGL(k_adam, dim3((kGParams + 255) / 256), dim3(256), 0, s, A);
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
int guide_train_rows(rx_ctx* c, uint32_t cap) {
  GuideState& G = *c->guide;
  hipLaunchKernelGGL(k_guide_encode, dim3(std::min<uint32_t>((cap + 255) / 256, 4096u)), dim3(256), 0, c->stream,
                     G.net(), G.pos.as<float4>(), G.dir.as<float4>(), G.rows.as<uint32_t>(), cap, G.X.as<uint16_t>());
  if (int r = hip_err(c, hipGetLastError(), "encode")) return r;
  return guide_train_device(c, G, cap, nullptr);
}

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
  if (int r = hip_err(c, G.rows.upload(&n, 4, s), "upload rows")) return r;
  hipLaunchKernelGGL(k_guide_encode, dim3(std::min<uint32_t>((n + 255) / 256, 4096u)), dim3(256), 0, s, G.net(),
                     G.pos.as<float4>(), G.dir.as<float4>(), G.rows.as<uint32_t>(), n, G.X.as<uint16_t>());
  return hip_err(c, hipGetLastError(), "encode");
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
    std::vector<float> lp((n + 255) / 256);
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
