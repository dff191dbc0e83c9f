// Oracle for the path-guiding network (no counterpart in the reference):
// FP64 frequency encoding, MLP 96-64-64-64-48 with ReLU, the 8-lobe
// von Mises-Fisher mixture + selection probability head, the one-sample-MIS
// KL loss, its analytic gradient, backpropagation and one Adam step.
//
// TEST INFRASTRUCTURE ONLY: checks the HIP MFMA kernels of renderx_amd.
//
// Output layout of the last layer (48 columns, 41 used):
//   lobe j (0..7): z[5j] weight logit, z[5j+1..5j+3] mean direction (unnormalised),
//                  z[5j+4] log concentration;  z[40] selection logit.
//   pi = softmax(logits), mu_j = m_j/|m_j|, kappa_j = exp(clamp(lk_j, -3, 8)),
//   alpha = 0.05 + 0.45 sigmoid(z[40]) (at most half of the samples guided).
//   vMF_j(w) = kappa/(2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))
//   q(w) = sum_j pi_j vMF_j(w);  p_mix = alpha q + (1 - alpha) p_bsdf
//   loss = -weight * log p_mix(w)
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

const int IN = 96, H = 64, OUT = 48, NL = 8;
const double PI = 3.14159265358979323846;

struct Head {
  double pi[NL], mu[NL][3], mlen[NL], kappa[NL], kclamped[NL], alpha, sig;
};

void head_of(const double* z, Head& h) {
  double mx = -1e300;
  for (int j = 0; j < NL; j++) mx = std::max(mx, z[5 * j]);
  double s = 0;
  for (int j = 0; j < NL; j++) {
    h.pi[j] = std::exp(z[5 * j] - mx);
    s += h.pi[j];
  }
  for (int j = 0; j < NL; j++) {
    h.pi[j] /= s;
    double m[3] = {z[5 * j + 1], z[5 * j + 2], z[5 * j + 3]};
    double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    l = std::max(l, 1e-6);
    h.mlen[j] = l;
    for (int k = 0; k < 3; k++) h.mu[j][k] = m[k] / l;
    double lk = z[5 * j + 4];
    h.kclamped[j] = (lk < -3.0 || lk > 8.0) ? 1.0 : 0.0;
    lk = std::min(8.0, std::max(-3.0, lk));
    h.kappa[j] = std::exp(lk);
  }
  h.sig = 1.0 / (1.0 + std::exp(-z[40]));
  h.alpha = 0.05 + 0.45 * h.sig;
}

double vmf(double kappa, const double* mu, const double* w) {
  double c = mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2];
  return kappa / (2 * PI * (1 - std::exp(-2 * kappa))) * std::exp(kappa * (c - 1));
}

}  // namespace

extern "C" {

// Frequency encoding: position normalised to [lo, hi] (12 frequencies) and the
// direction (4 frequencies), sin/cos of 2^k pi x per component.
void orc_mlp_encode(int n, const float* pos, const float* dir, const float* lo, const float* hi, double* out) {
  for (int r = 0; r < n; r++) {
    double* o = out + (size_t)r * IN;
    for (int a = 0; a < 3; a++) {
      double p = (pos[3 * r + a] - lo[a]) / (double)(hi[a] - lo[a]);
      p = std::min(1.0, std::max(0.0, p));
      for (int k = 0; k < 12; k++) {
        double f = std::ldexp(PI, k) * p;
        o[6 * k + 2 * a] = std::sin(f);
        o[6 * k + 2 * a + 1] = std::cos(f);
      }
      double d = dir[3 * r + a];
      for (int k = 0; k < 4; k++) {
        double f = std::ldexp(PI, k) * d;
        o[72 + 6 * k + 2 * a] = std::sin(f);
        o[72 + 6 * k + 2 * a + 1] = std::cos(f);
      }
    }
  }
}

// Forward pass. W: W1[96x64], W2[64x64], W3[64x64], W4[64x48] row-major
// (in x out), concatenated; b: 64+64+64+48.  acts (optional): a1,a2,a3 (n x 64 each).
void orc_mlp_forward(int n, const double* x, const double* W, const double* b, double* z4, double* acts) {
  const double* W1 = W;
  const double* W2 = W1 + IN * H;
  const double* W3 = W2 + H * H;
  const double* W4 = W3 + H * H;
  const double *b1 = b, *b2 = b + H, *b3 = b + 2 * H, *b4 = b + 3 * H;
  for (int r = 0; r < n; r++) {
    double a1[H], a2[H], a3[H];
    const double* xr = x + (size_t)r * IN;
    for (int j = 0; j < H; j++) {
      double s = b1[j];
      for (int i = 0; i < IN; i++) s += xr[i] * W1[i * H + j];
      a1[j] = s > 0 ? s : 0;
    }
    for (int j = 0; j < H; j++) {
      double s = b2[j];
      for (int i = 0; i < H; i++) s += a1[i] * W2[i * H + j];
      a2[j] = s > 0 ? s : 0;
    }
    for (int j = 0; j < H; j++) {
      double s = b3[j];
      for (int i = 0; i < H; i++) s += a2[i] * W3[i * H + j];
      a3[j] = s > 0 ? s : 0;
    }
    for (int j = 0; j < OUT; j++) {
      double s = b4[j];
      for (int i = 0; i < H; i++) s += a3[i] * W4[i * OUT + j];
      z4[(size_t)r * OUT + j] = s;
    }
    if (acts) {
      memcpy(acts + ((size_t)r * 3 + 0) * H, a1, sizeof a1);
      memcpy(acts + ((size_t)r * 3 + 1) * H, a2, sizeof a2);
      memcpy(acts + ((size_t)r * 3 + 2) * H, a3, sizeof a3);
    }
  }
}

// Mixture pdf q(w), selection probability alpha, and the loss with its
// gradient with respect to z4 (per row).  w: sampled direction, pb: BSDF
// density at w, wt: sample weight F/p_sample.
double orc_mlp_loss(int n, const double* z4, const float* w, const float* pb, const float* wt, double* dz4,
                    double* q_out, double* alpha_out) {
  double total = 0;
  for (int r = 0; r < n; r++) {
    Head h;
    head_of(z4 + (size_t)r * OUT, h);
    const double om[3] = {w[3 * r], w[3 * r + 1], w[3 * r + 2]};
    double v[NL], q = 0;
    for (int j = 0; j < NL; j++) {
      v[j] = vmf(h.kappa[j], h.mu[j], om);
      q += h.pi[j] * v[j];
    }
    double pm = h.alpha * q + (1 - h.alpha) * pb[r];
    pm = std::max(pm, 1e-12);
    total += -wt[r] * std::log(pm);
    if (q_out) q_out[r] = q;
    if (alpha_out) alpha_out[r] = h.alpha;
    if (!dz4) continue;
    double* g = dz4 + (size_t)r * OUT;
    for (int c = 0; c < OUT; c++) g[c] = 0;
    double gp = -wt[r] / pm;
    double dq = gp * h.alpha;
    g[40] = gp * (q - pb[r]) * 0.45 * h.sig * (1 - h.sig);
    for (int j = 0; j < NL; j++) {
      g[5 * j] = dq * h.pi[j] * (v[j] - q);
      double dv = dq * h.pi[j];
      double k = h.kappa[j];
      double c = h.mu[j][0] * om[0] + h.mu[j][1] * om[1] + h.mu[j][2] * om[2];
      double e = std::exp(-2 * k);
      double dlogv_dk = 1.0 / k - 2 * e / (1 - e) + (c - 1);
      g[5 * j + 4] = h.kclamped[j] ? 0.0 : dv * v[j] * dlogv_dk * k;
      double s = dv * v[j] * k / h.mlen[j];
      for (int a = 0; a < 3; a++) g[5 * j + 1 + a] = s * (om[a] - c * h.mu[j][a]);
    }
  }
  return total;
}

// Backward: dW (same layout as W) and db summed over the rows.
void orc_mlp_backward(int n, const double* x, const double* W, const double* acts, const double* dz4, double* dW,
                      double* db) {
  const double* W2 = W + IN * H;
  const double* W3 = W2 + H * H;
  const double* W4 = W3 + H * H;
  double* dW1 = dW;
  double* dW2 = dW1 + IN * H;
  double* dW3 = dW2 + H * H;
  double* dW4 = dW3 + H * H;
  memset(dW, 0, sizeof(double) * (IN * H + 2 * H * H + H * OUT));
  memset(db, 0, sizeof(double) * (3 * H + OUT));
  for (int r = 0; r < n; r++) {
    const double* a1 = acts + ((size_t)r * 3 + 0) * H;
    const double* a2 = acts + ((size_t)r * 3 + 1) * H;
    const double* a3 = acts + ((size_t)r * 3 + 2) * H;
    const double* d4 = dz4 + (size_t)r * OUT;
    const double* xr = x + (size_t)r * IN;
    double d3[H], d2[H], d1[H];
    for (int i = 0; i < H; i++) {
      double s = 0;
      for (int j = 0; j < OUT; j++) s += d4[j] * W4[i * OUT + j];
      d3[i] = a3[i] > 0 ? s : 0;
    }
    for (int i = 0; i < H; i++) {
      double s = 0;
      for (int j = 0; j < H; j++) s += d3[j] * W3[i * H + j];
      d2[i] = a2[i] > 0 ? s : 0;
    }
    for (int i = 0; i < H; i++) {
      double s = 0;
      for (int j = 0; j < H; j++) s += d2[j] * W2[i * H + j];
      d1[i] = a1[i] > 0 ? s : 0;
    }
    for (int i = 0; i < H; i++)
      for (int j = 0; j < OUT; j++) dW4[i * OUT + j] += a3[i] * d4[j];
    for (int i = 0; i < H; i++)
      for (int j = 0; j < H; j++) {
        dW3[i * H + j] += a2[i] * d3[j];
        dW2[i * H + j] += a1[i] * d2[j];
      }
    for (int i = 0; i < IN; i++)
      for (int j = 0; j < H; j++) dW1[i * H + j] += xr[i] * d1[j];
    for (int j = 0; j < H; j++) {
      db[j] += d1[j];
      db[H + j] += d2[j];
      db[2 * H + j] += d3[j];
    }
    for (int j = 0; j < OUT; j++) db[3 * H + j] += d4[j];
  }
}

// One Adam step (Kingma & Ba 2015) on np parameters at step t (1-based).
void orc_adam(int np, double* p, double* m, double* v, const double* g, double lr, double b1, double b2, double eps,
              int t) {
  double c1 = 1 - std::pow(b1, t), c2 = 1 - std::pow(b2, t);
  for (int i = 0; i < np; i++) {
    m[i] = b1 * m[i] + (1 - b1) * g[i];
    v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
    p[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
  }
}

}  // extern "C"
