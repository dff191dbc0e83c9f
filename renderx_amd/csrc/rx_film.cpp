// Film I/O and image comparison (host only): the reference's PFM writer and
// reader and the RMSE of its cmp tool, plus relMSE for the equal-time study.
//
// Layout convention of this library: RGB float32, row-major y*W+x with y = 0
// the top row (rx_render's output, MCRenderer::renderPixels' vector<vec3f>).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rxhip.h"

namespace {

struct File {
  FILE* f;
  explicit File(FILE* x) : f(x) {}
  ~File() {
    if (f) fclose(f);
  }
};

float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Skips whitespace and reads one header token.
bool token(FILE* f, std::string& out) {
  out.clear();
  int ch;
  while ((ch = fgetc(f)) != EOF && (ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t')) {
  }
  if (ch == EOF) return false;
  do {
    out.push_back((char)ch);
  } while ((ch = fgetc(f)) != EOF && !(ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t'));
  // the single whitespace byte after the scale token separates it from the data
  return true;
}

}  // namespace

extern "C" {

// MCRenderer::saveImagePFM (src/MCRenderer.cpp:556-574): header
// "PF\n%d %d\n-1.000000\n" (negative scale = little endian), then rows from
// the bottom of the image to the top, R G B float32 per pixel.
int rx_save_pfm(const char* path, int32_t width, int32_t height, const float* rgb) {
  if (!path || !rgb || width <= 0 || height <= 0) return RX_ERR_INVALID;
  File f(fopen(path, "wb"));
  if (!f.f) return RX_ERR_INVALID;
  if (fprintf(f.f, "PF\n%d %d\n-1.000000\n", width, height) < 0) return RX_ERR_INVALID;
  for (int32_t j = height - 1; j >= 0; j--)
    if (fwrite(rgb + 3 * (size_t)j * width, sizeof(float), 3 * (size_t)width, f.f) != 3 * (size_t)width)
      return RX_ERR_INVALID;
  return RX_OK;
}

// MCRenderer::showCurrentResult's time-stamped copy (src/MCRenderer.cpp:465-490):
// save_path with its extension replaced by "_time_<time>_iter_<iter>.pfm".
int rx_save_pfm_snapshot(const char* save_path, uint32_t time, uint32_t iter, int32_t width, int32_t height,
                         const float* rgb) {
  if (!save_path) return RX_ERR_INVALID;
  std::string name(save_path);
  for (size_t i = name.size(); i-- > 0;) {  // drop the extension with its dot (the whole name if there is none)
    const char ch = name[i];
    name.pop_back();
    if (ch == '.') break;
  }
  name += "_time_" + std::to_string(time) + "_iter_" + std::to_string(iter) + ".pfm";
  return rx_save_pfm(name.c_str(), width, height, rgb);
}

// readImagePFM (include/visualizeUtils.h:23-48) with the header parsed as the
// PFM format defines it ("PF", width, height, scale; a negative scale means
// little endian; one whitespace byte before the data).  The reference parses
// the scale with %d, which leaves its own writer's ".000000\n" in front of the
// data; this reader does not reproduce that shift.  clamp != 0 clamps every
// channel to [0, 4] like the reference.  rgb may be NULL to query the size.
int rx_read_pfm(const char* path, int32_t* width, int32_t* height, float* rgb, int32_t clamp) {
  if (!path || !width || !height) return RX_ERR_INVALID;
  File f(fopen(path, "rb"));
  if (!f.f) return RX_ERR_INVALID;
  std::string t;
  if (!token(f.f, t) || t != "PF") return RX_ERR_INVALID;
  std::string ws, hs, ss;
  if (!token(f.f, ws) || !token(f.f, hs) || !token(f.f, ss)) return RX_ERR_INVALID;
  char* e = nullptr;
  const long w = strtol(ws.c_str(), &e, 10);
  if (*e || w <= 0) return RX_ERR_INVALID;
  const long h = strtol(hs.c_str(), &e, 10);
  if (*e || h <= 0 || (uint64_t)w * (uint64_t)h > (1ull << 30)) return RX_ERR_INVALID;
  const double scale = strtod(ss.c_str(), &e);
  if (*e || scale == 0.0) return RX_ERR_INVALID;
  *width = (int32_t)w;
  *height = (int32_t)h;
  if (!rgb) return RX_OK;
  const bool big = scale > 0;
  std::vector<float> row(3 * (size_t)w);
  for (long j = h - 1; j >= 0; j--) {
    if (fread(row.data(), sizeof(float), row.size(), f.f) != row.size()) return RX_ERR_INVALID;
    for (size_t k = 0; k < row.size(); k++) {
      float v = row[k];
      if (big) {
        uint32_t u;
        memcpy(&u, &v, 4);
        u = (u >> 24) | ((u >> 8) & 0xff00u) | ((u << 8) & 0xff0000u) | (u << 24);
        memcpy(&v, &u, 4);
      }
      rgb[3 * (size_t)j * w + k] = clamp ? clampf(v, 0.0f, 4.0f) : v;
    }
  }
  return RX_OK;
}

// CalculateRMSE (cmp/compare.cpp:85-124): over the pixels whose mask is set
// (mask NULL = all), sqrt(sum |ref - img|^2 / count) with the difference and
// its length in float (vec3f, channels in the reference image's B, G, R
// order) and the square and sum in double.
int rx_image_rmse(int32_t width, int32_t height, const float* img, const float* ref, const uint8_t* mask,
                  double* rmse) {
  if (!img || !ref || !rmse || width <= 0 || height <= 0) return RX_ERR_INVALID;
  double total = 0.0;
  uint64_t count = 0;
  for (int32_t x = 0; x < width; x++) {
    for (int32_t y = 0; y < height; y++) {
      const size_t i = (size_t)y * width + x;
      if (mask && !mask[i]) continue;
      const float db = ref[3 * i + 2] - img[3 * i + 2];
      const float dg = ref[3 * i + 1] - img[3 * i + 1];
      const float dr = ref[3 * i] - img[3 * i];
      float d = 0.0f;
      d += db * db;
      d += dg * dg;
      d += dr * dr;
      const double len = (double)sqrtf(d);
      total += len * len;
      count++;
    }
  }
  *rmse = count ? std::sqrt(total / (double)count) : 0.0;
  return RX_OK;
}

// relMSE: mean over pixels and channels of (img - ref)^2 / (ref^2 + eps), in double.
int rx_image_relmse(int32_t width, int32_t height, const float* img, const float* ref, float eps, double* relmse) {
  if (!img || !ref || !relmse || width <= 0 || height <= 0 || !(eps > 0)) return RX_ERR_INVALID;
  const size_t n = 3 * (size_t)width * height;
  double total = 0.0;
  for (size_t i = 0; i < n; i++) {
    const double d = (double)img[i] - (double)ref[i];
    total += d * d / ((double)ref[i] * ref[i] + eps);
  }
  *relmse = total / (double)n;
  return RX_OK;
}

}  // extern "C"
