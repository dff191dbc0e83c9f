// rx_cmp: the reference's cmp tool (cmp/compare.cpp:242-331, main :333-343)
// on Linux.  For every *.pfm in <input_dir> whose name carries "_time_<T>"
// and "_iter_<I>" tokens, the RMSE against <ref_img> (both read with the
// reference's [0, 4] channel clamp, optional mask of 0/1 integers listed from
// the bottom row up); output lines "<time> <rmse>" sorted by time, like the
// reference.  With --relmse a third column holds the relMSE (eps 1e-2) of the
// unclamped images.
//
//   rx_cmp [--relmse] <input_dir> <ref_img> <output_file> [<mask_file>]
#include <dirent.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rxhip.h"

extern "C" {
int rx_read_pfm(const char* path, int32_t* width, int32_t* height, float* rgb, int32_t clamp);
int rx_image_rmse(int32_t width, int32_t height, const float* img, const float* ref, const uint8_t* mask,
                  double* rmse);
int rx_image_relmse(int32_t width, int32_t height, const float* img, const float* ref, float eps, double* relmse);
}

namespace {

struct Row {
  int iter, time;
  double rmse, relmse;
};

bool read(const std::string& path, int clamp, int& w, int& h, std::vector<float>& rgb) {
  int32_t ww = 0, hh = 0;
  if (rx_read_pfm(path.c_str(), &ww, &hh, nullptr, clamp) != RX_OK) return false;
  rgb.resize(3 * (size_t)ww * hh);
  if (rx_read_pfm(path.c_str(), &ww, &hh, rgb.data(), clamp) != RX_OK) return false;
  w = ww;
  h = hh;
  return true;
}

// "<...>_time_<T>_iter_<I>.pfm" tokens split at '_' and '.' (cmp/compare.cpp:285-310)
void parse_name(const std::string& name, int& iter, int& time) {
  iter = time = -1;
  bool iter_next = false, time_next = false;
  std::string tok;
  for (size_t i = 0; i <= name.size(); i++) {
    if (i == name.size() || name[i] == '_' || name[i] == '.') {
      if (iter_next) iter = atoi(tok.c_str());
      if (time_next) time = atoi(tok.c_str());
      iter_next = tok == "iter";
      time_next = tok == "time";
      tok.clear();
      continue;
    }
    tok += name[i];
  }
}

}  // namespace

int main(int argc, char** argv) {
  bool relmse = false;
  std::vector<std::string> args;
  for (int i = 1; i < argc; i++) {
    if (strcmp(argv[i], "--relmse") == 0) relmse = true;
    else args.push_back(argv[i]);
  }
  if (args.size() < 3) {
    fprintf(stderr, "usage: rx_cmp [--relmse] <input_dir> <ref_img> <output_file> [<mask_file>]\n");
    return 2;
  }
  std::string root = args[0];
  if (!root.empty() && root.back() != '/') root += '/';
  int w = 0, h = 0;
  std::vector<float> ref, ref_raw;
  if (!read(args[1], 1, w, h, ref) || !read(args[1], 0, w, h, ref_raw)) {
    fprintf(stderr, "rx_cmp: cannot read %s\n", args[1].c_str());
    return 1;
  }
  std::vector<uint8_t> mask((size_t)w * h, 1);
  if (args.size() >= 4) {
    FILE* fm = fopen(args[3].c_str(), "r");
    if (!fm) {
      fprintf(stderr, "rx_cmp: cannot read %s\n", args[3].c_str());
      return 1;
    }
    for (int y = h - 1; y >= 0; y--)
      for (int x = 0; x < w; x++) {
        int f = 0;
        if (fscanf(fm, "%d ", &f) != 1) f = 0;
        mask[(size_t)y * w + x] = (f == 1);
      }
    fclose(fm);
  }
  DIR* dir = opendir(root.c_str());
  if (!dir) {
    fprintf(stderr, "rx_cmp: cannot open %s\n", root.c_str());
    return 1;
  }
  std::vector<Row> rows;
  while (dirent* e = readdir(dir)) {
    const std::string name = e->d_name;
    if (name.size() < 4 || name.compare(name.size() - 4, 4, ".pfm") != 0) continue;
    printf("%s\n", name.c_str());
    int iw = 0, ih = 0;
    std::vector<float> img, img_raw;
    if (!read(root + name, 1, iw, ih, img) || iw != w || ih != h) {
      fprintf(stderr, "rx_cmp: skipping %s (unreadable or another size)\n", name.c_str());
      continue;
    }
    Row r;
    parse_name(name, r.iter, r.time);
    if (r.iter < 0 || r.time < 0) continue;
    rx_image_rmse(w, h, img.data(), ref.data(), mask.data(), &r.rmse);
    r.relmse = 0.0;
    if (relmse && read(root + name, 0, iw, ih, img_raw))
      rx_image_relmse(w, h, img_raw.data(), ref_raw.data(), 1e-2f, &r.relmse);
    rows.push_back(r);
  }
  closedir(dir);
  std::stable_sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) { return a.time < b.time; });
  FILE* fp = fopen(args[2].c_str(), "w");
  if (!fp) {
    fprintf(stderr, "rx_cmp: cannot write %s\n", args[2].c_str());
    return 1;
  }
  for (const Row& r : rows) {
    if (relmse) fprintf(fp, "%d %.8lf %.8lf\n", r.time, r.rmse, r.relmse);
    else fprintf(fp, "%d %.8lf\n", r.time, r.rmse);
  }
  fclose(fp);
  return 0;
}
