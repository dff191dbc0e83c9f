// Scene ingestion (host only): ConfigManager::load (src/ConfigManager.cpp:166-405)
// for the object types of this path, with SimpleShape::loadShape's OBJ/SHP
// reading (src/SimpleShape.cpp:44-313), the XML helpers readVec / readMatrix
// (src/xmlHelper.cpp) and the world transform getWorldVertexPosition
// (include/SimpleShape.h:44).  The result is the arrays rx_load_scene takes,
// the camera and the render parameters.
//
// Face normals are the path's shading normals: OBJ vertex normals and texture
// coordinates are parsed and dropped, textures are rejected.
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rxhip.h"

namespace {

// ------------------------------------------------------------------ XML
struct XNode {
  std::string tag, text;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<XNode>> kids;
  const XNode* first(const std::string& t) const {
    for (auto& k : kids)
      if (k->tag == t) return k.get();
    return nullptr;
  }
  const char* attr(const std::string& a) const {
    for (auto& kv : attrs)
      if (kv.first == a) return kv.second.c_str();
    return nullptr;
  }
  const char* value() const { return text.c_str(); }
};

struct XParser {
  const std::string& s;
  size_t i = 0;
  std::string err;
  explicit XParser(const std::string& src) : s(src) {}
  void ws() {
    while (i < s.size() && isspace((unsigned char)s[i])) i++;
  }
  static std::string unescape(const std::string& t) {
    std::string o;
    for (size_t k = 0; k < t.size(); k++) {
      if (t[k] == '&') {
        static const char* names[] = {"&lt;", "&gt;", "&amp;", "&quot;", "&apos;"};
        static const char subs[] = {'<', '>', '&', '"', '\''};
        bool done = false;
        for (int n = 0; n < 5 && !done; n++)
          if (t.compare(k, strlen(names[n]), names[n]) == 0) {
            o += subs[n];
            k += strlen(names[n]) - 1;
            done = true;
          }
        if (!done) o += '&';
      } else {
        o += t[k];
      }
    }
    return o;
  }
  bool skip_misc() {  // declarations, comments, doctype
    for (;;) {
      ws();
      if (s.compare(i, 4, "<!--") == 0) {
        size_t e = s.find("-->", i + 4);
        if (e == std::string::npos) return false;
        i = e + 3;
      } else if (s.compare(i, 2, "<?") == 0) {
        size_t e = s.find("?>", i + 2);
        if (e == std::string::npos) return false;
        i = e + 2;
      } else if (s.compare(i, 2, "<!") == 0) {
        size_t e = s.find('>', i + 2);
        if (e == std::string::npos) return false;
        i = e + 1;
      } else {
        return true;
      }
    }
  }
  std::unique_ptr<XNode> element() {
    if (i >= s.size() || s[i] != '<') return nullptr;
    i++;
    auto n = std::make_unique<XNode>();
    while (i < s.size() && !isspace((unsigned char)s[i]) && s[i] != '>' && s[i] != '/') n->tag += s[i++];
    for (;;) {
      ws();
      if (i >= s.size()) return nullptr;
      if (s[i] == '/') {
        if (s.compare(i, 2, "/>") != 0) return nullptr;
        i += 2;
        return n;
      }
      if (s[i] == '>') {
        i++;
        break;
      }
      std::string a;
      while (i < s.size() && s[i] != '=' && !isspace((unsigned char)s[i])) a += s[i++];
      ws();
      if (i >= s.size() || s[i] != '=') return nullptr;
      i++;
      ws();
      if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) return nullptr;
      const char q = s[i++];
      size_t e = s.find(q, i);
      if (e == std::string::npos) return nullptr;
      n->attrs.emplace_back(a, unescape(s.substr(i, e - i)));
      i = e + 1;
    }
    bool have_text = false;
    for (;;) {
      if (i >= s.size()) return nullptr;
      if (s.compare(i, 4, "<!--") == 0) {
        size_t e = s.find("-->", i + 4);
        if (e == std::string::npos) return nullptr;
        i = e + 3;
      } else if (s.compare(i, 2, "</") == 0) {
        size_t e = s.find('>', i);
        if (e == std::string::npos) return nullptr;
        i = e + 1;
        return n;
      } else if (s[i] == '<') {
        auto k = element();
        if (!k) return nullptr;
        n->kids.push_back(std::move(k));
      } else {
        size_t e = s.find('<', i);
        if (e == std::string::npos) return nullptr;
        std::string t = s.substr(i, e - i);
        bool blank = true;
        for (char ch : t) blank = blank && isspace((unsigned char)ch);
        if (!blank && !have_text) {  // rapidxml's value(): the element's first data node
          n->text = unescape(t);
          have_text = true;
        }
        i = e;
      }
    }
  }
  // the document's top-level elements (rapidxml's doc->first_node siblings)
  bool parse(std::vector<std::unique_ptr<XNode>>& tops) {
    for (;;) {
      if (!skip_misc()) return false;
      if (i >= s.size()) return true;
      auto n = element();
      if (!n) return false;
      tops.push_back(std::move(n));
    }
  }
};

// ------------------------------------------------------------------ math (float, the reference's order)
struct M4 {
  float e[4][4];  // e[row][col]
};
M4 identity() {
  M4 m;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) m.e[r][c] = r == c ? 1.0f : 0.0f;
  return m;
}
// readMatrix (src/xmlHelper.cpp:4-17): "[a, b, c, d][...]..." rows as written
bool read_matrix(const char* v, M4& m) {
  float a[16];
  if (sscanf(v, " [%f, %f, %f, %f][%f, %f, %f, %f][%f, %f, %f, %f][%f, %f, %f, %f]", a, a + 1, a + 2, a + 3, a + 4,
             a + 5, a + 6, a + 7, a + 8, a + 9, a + 10, a + 11, a + 12, a + 13, a + 14, a + 15) != 16)
    return false;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) m.e[r][c] = a[4 * r + c];
  return true;
}
// readVec (src/xmlHelper.cpp:19-24): "(x, y, z)"
bool read_vec(const char* v, float* out) { return sscanf(v, " (%f, %f, %f)", out, out + 1, out + 2) == 3; }
// matrix4 * vec4 (include/nvMatrix.h:415-423): r[i] = v0 m(i,0) + v1 m(i,1) + v2 m(i,2) + v3 m(i,3)
void xform(const M4& m, const float* v, float* out) {
  for (int r = 0; r < 3; r++) {
    float x = v[0] * m.e[r][0];
    x = x + v[1] * m.e[r][1];
    x = x + v[2] * m.e[r][2];
    x = x + 1.0f * m.e[r][3];
    out[r] = x;
  }
}

// ------------------------------------------------------------------ shapes
struct Shape {
  std::string name;
  std::vector<float> v;       // 3 per vertex (object space)
  std::vector<uint32_t> f;    // 3 per triangle
  M4 transform = identity();
};

// SimpleShape::loadShape (src/SimpleShape.cpp:44-313).  split: one shape per
// "g" group (triangles only; quads are not split out, like the reference);
// normalize: unitize the whole file and bake that into the split shapes.
bool load_shape(const std::string& path, bool normalize, bool split, Shape& whole, std::vector<Shape>& parts,
                std::string& err) {
  if (path.size() < 4) {
    err = "shape file name too short: " + path;
    return false;
  }
  const std::string suffix = path.substr(path.size() - 4);
  if (suffix == ".shp") {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
      err = "cannot open " + path;
      return false;
    }
    int n = 0;
    bool ok = fread(&n, 4, 1, f) == 1 && n >= 0;
    if (ok) {
      whole.v.resize(n);
      ok = fread(whole.v.data(), 4, n, f) == (size_t)n;
    }
    if (ok) ok = fread(&n, 4, 1, f) == 1 && n >= 0;
    if (ok) {
      whole.f.resize(n);
      ok = fread(whole.f.data(), 4, n, f) == (size_t)n;
    }
    fclose(f);
    if (!ok || whole.v.size() % 3 || whole.f.size() % 3) {
      err = "bad shp file " + path;
      return false;
    }
  } else if (suffix == ".obj") {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) {
      err = "cannot open " + path;
      return false;
    }
    char line[4096], attrib[4096], p[4][4096];
    Shape* cur = nullptr;
    uint32_t cur_fi = 0;
    while (fgets(line, sizeof line, f)) {
      if (line[0] == '#') continue;
      attrib[0] = 0;
      const int num = sscanf(line, "%s %s %s %s", attrib, p[0], p[1], p[2]);
      const int num_rect = sscanf(line, "%s %s %s %s %s", attrib, p[0], p[1], p[2], p[3]);
      if (strcmp(attrib, "g") == 0 && split) {
        parts.emplace_back();
        cur = &parts.back();
        cur->name = num >= 2 ? p[0] : "";
        cur_fi = (uint32_t)(whole.v.size() / 3);
      }
      auto vidx = [](const char* s) {  // "v", "v/t", "v/t/n", "v//n": the vertex index
        int a = 0;
        sscanf(s, "%d", &a);
        return a - 1;
      };
      if (strcmp(attrib, "f") == 0 && num_rect == 5) {  // quad: (0,1,2) and (0,2,3)
        const int t[4] = {vidx(p[0]), vidx(p[1]), vidx(p[2]), vidx(p[3])};
        const int q[6] = {t[0], t[1], t[2], t[0], t[2], t[3]};
        for (int k = 0; k < 6; k++) whole.f.push_back((uint32_t)q[k]);
        continue;
      }
      if (num != 4) continue;
      if (strcmp(attrib, "v") == 0) {
        float x = 0, y = 0, z = 0;
        sscanf(p[0], "%f", &x);
        sscanf(p[1], "%f", &y);
        sscanf(p[2], "%f", &z);
        whole.v.insert(whole.v.end(), {x, y, z});
        if (split && cur) cur->v.insert(cur->v.end(), {x, y, z});
      } else if (strcmp(attrib, "f") == 0) {
        const int t[3] = {vidx(p[0]), vidx(p[1]), vidx(p[2])};
        for (int k = 0; k < 3; k++) whole.f.push_back((uint32_t)t[k]);
        if (split && cur)
          for (int k = 0; k < 3; k++) cur->f.push_back((uint32_t)t[k] - cur_fi);
      }
    }
    fclose(f);
  } else {
    err = "unknown shape format " + path;
    return false;
  }
  const size_t nv = whole.v.size() / 3;
  for (uint32_t i : whole.f)
    if (i >= nv) {
      err = "face index out of range in " + path;
      return false;
    }
  for (auto& s : parts)
    for (uint32_t i : s.f)
      if (i >= s.v.size() / 3) {
        err = "group face index out of range in " + path;
        return false;
      }
  if (normalize && nv) {  // SimpleShape::unitize (src/SimpleShape.cpp:350-364), baked into the parts
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = whole.v[a];
    for (size_t k = 0; k < nv; k++)
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], whole.v[3 * k + a]);
        hi[a] = std::max(hi[a], whole.v[3 * k + a]);
      }
    float center[3], delta[3];
    for (int a = 0; a < 3; a++) {
      center[a] = lo[a] + hi[a];
      center[a] = center[a] / 2;
      delta[a] = hi[a] - lo[a];
    }
    float maxLen = std::max(delta[0], delta[1]);
    maxLen = std::max(maxLen, delta[2]);
    M4 u = identity();
    for (int a = 0; a < 3; a++) {
      u.e[a][a] = 2 / maxLen;
      u.e[a][3] = -2 / maxLen * center[a];
    }
    for (size_t k = 0; k < nv; k++)
      for (int a = 0; a < 3; a++) {
        float x = whole.v[3 * k + a] - center[a];
        const float h = maxLen / 2;
        if (h != 0) x = x / h;
        whole.v[3 * k + a] = x;
      }
    for (auto& s : parts) {  // applyTransform with the unitize matrix
      std::vector<float> o(s.v.size());
      for (size_t k = 0; k < s.v.size() / 3; k++) xform(u, &s.v[3 * k], &o[3 * k]);
      s.v.swap(o);
    }
  }
  return true;
}

// ------------------------------------------------------------------ config
struct Loader {
  std::string root, current_path, err;
  std::map<std::string, std::vector<std::unique_ptr<XNode>>> docs;
  std::vector<float> verts;
  std::vector<uint32_t> tris;
  std::vector<rx_object_desc> objs;
  rx_camera_desc cam{};
  rx_render_params prm{};
  std::string save_path;
  double runtime = 0.0;

  static bool exists(const std::string& p) { return access(p.c_str(), F_OK) == 0; }
  static std::string folder(const std::string& s) {  // ConfigManager::getFolder
    size_t a = s.rfind('/'), b = s.rfind('\\');
    size_t pos = std::min(a == std::string::npos ? s.size() : a, b == std::string::npos ? s.size() : b);
    return pos < s.size() ? s.substr(0, pos) : "";
  }
  std::string resolve(const std::string& p) const { return exists(p) ? p : root + "/" + p; }

  const std::vector<std::unique_ptr<XNode>>* doc(const std::string& path) {
    const std::string full = resolve(path);
    auto it = docs.find(full);
    if (it != docs.end()) return &it->second;
    FILE* f = fopen(full.c_str(), "rb");
    if (!f) return nullptr;
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    fclose(f);
    std::vector<std::unique_ptr<XNode>> tops;
    XParser px(text);
    if (!px.parse(tops)) {
      err = "cannot parse XML " + full;
      return nullptr;
    }
    return &(docs[full] = std::move(tops));
  }
  static const XNode* pick(const XNode* n, const std::string& name) {
    if (!n->attr("Name") || name.empty()) return n;
    if (n->attr("Name") && name == n->attr("Name")) return n;
    return nullptr;
  }
  // ConfigManager::findNode (src/ConfigManager.cpp:407-446)
  const XNode* find(const std::string& path, const std::string& tag, const std::string& name) {
    if (path.empty()) return nullptr;
    auto* d = doc(path);
    if (!d) return nullptr;
    for (auto& n : *d)
      if (n->tag == tag)
        if (const XNode* r = pick(n.get(), name)) return r;
    return nullptr;
  }
  static const XNode* find(const XNode* root, const std::string& tag, const std::string& name) {
    for (auto& n : root->kids)
      if (n->tag == tag)
        if (const XNode* r = pick(n.get(), name)) return r;
    return nullptr;
  }
  std::pair<std::string, std::string> path_name(const XNode* n) const {  // getPathAndName
    return {n->first("path") ? n->first("path")->value() : current_path, n->first("name") ? n->first("name")->value() : ""};
  }
  const XNode* material(const XNode* obj, const XNode* scene) {
    const XNode* m = obj->first("Material");
    if (!m) return nullptr;
    auto pn = path_name(m);
    const XNode* r = find(pn.first, "Material", pn.second);
    if (!r && m->first("name")) r = find(scene, "Material", pn.second);
    return r ? r : m;
  }
  bool vec(const XNode* n, const char* tag, float* out, bool required = true) {
    const XNode* k = n ? n->first(tag) : nullptr;
    if (!k) {
      if (required) err = std::string("material lacks <") + tag + ">";
      return !required;
    }
    if (!read_vec(k->value(), out)) {
      err = std::string("bad vector in <") + tag + ">: " + k->value();
      return false;
    }
    return true;
  }
  bool num(const XNode* n, const char* tag, float& out, bool required = true) {
    const XNode* k = n ? n->first(tag) : nullptr;
    if (!k) {
      if (required) err = std::string("material lacks <") + tag + ">";
      return !required;
    }
    out = (float)atof(k->value());
    return true;
  }
  // ConfigManager::generateSceneObject (src/ConfigManager.cpp:24-164) for the path's types
  bool material_desc(const XNode* mat, rx_object_desc& d) {
    memset(&d, 0, sizeof(d));
    const XNode* ty = mat->first("type");
    if (!ty) {
      err = "material without <type>";
      return false;
    }
    const std::string t = ty->value();
    if (mat->first("BumpTexture") || mat->first("Texture") || mat->first("ColorTexture") || mat->first("CoeffTexture")) {
      err = "textures are not supported on this path (material " + t + ")";
      return false;
    }
    if (t == "Diffuse") {
      d.type = RX_OBJ_DIFFUSE;
      return vec(mat, "color", d.color);
    }
    if (t == "Glossy") {
      d.type = RX_OBJ_GLOSSY;
      return vec(mat, "color", d.color) && num(mat, "coeff", d.exponent);
    }
    if (t == "Phong") {
      d.type = RX_OBJ_PHONG;
      const XNode* g = mat->first("GlossyComponent");
      const XNode* df = mat->first("DiffuseComponent");
      if (!g || !df) {
        err = "Phong material needs GlossyComponent and DiffuseComponent";
        return false;
      }
      return vec(g, "color", d.color) && num(g, "coeff", d.exponent) && vec(df, "color", d.color2);
    }
    if (t == "Refractive") {
      d.type = RX_OBJ_REFRACTIVE;
      d.color[0] = d.color[1] = d.color[2] = 1.0f;
      d.color2[0] = d.color2[1] = d.color2[2] = 1.0f;
      d.ior = 1.0f;
      return num(mat, "coeff", d.ior, false) && vec(mat, "surfColor", d.color, false) &&
             vec(mat, "decayColor", d.color2, false);
    }
    if (t == "Emissive") {
      d.type = RX_OBJ_EMISSIVE;
      return vec(mat, "color", d.color);
    }
    if (t == "VPM") {
      d.type = RX_OBJ_VPM;
      return vec(mat, "sigma_s", d.color) && vec(mat, "sigma_a", d.color2) && num(mat, "inrate", d.ior) &&
             num(mat, "g", d.g);
    }
    err = "material type " + t + " is not on this path";
    return false;
  }
  void add_object(const Shape& s, const M4& world, const rx_object_desc& d) {
    const uint32_t base = (uint32_t)(verts.size() / 3);
    for (size_t k = 0; k < s.v.size() / 3; k++) {
      float w[3];
      xform(world, &s.v[3 * k], w);
      verts.insert(verts.end(), w, w + 3);
    }
    rx_object_desc o = d;
    o.tri_begin = (uint32_t)(tris.size() / 3);
    o.tri_count = (uint32_t)(s.f.size() / 3);
    for (uint32_t i : s.f) tris.push_back(base + i);
    objs.push_back(o);
  }

  bool load(const std::string& config) {
    root = folder(config);
    const XNode* cfg = find(config, "Config", "");
    if (!cfg) {
      if (err.empty()) err = "no <Config> in " + config;
      return false;
    }
    const XNode* sc_cfg = cfg->first("Scene");
    if (!sc_cfg) {
      err = "<Config> has no <Scene>";
      return false;
    }
    auto pn = path_name(sc_cfg);
    const XNode* scene = find(pn.first, "Scene", pn.second);
    if (!scene) scene = find(config, "Scene", pn.second);
    if (!scene) {
      if (err.empty()) err = "scene not found";
      return false;
    }
    current_path = pn.first;
    for (auto& g : scene->kids) {
      if (g->tag != "GroupObject") continue;
      const XNode* fp = g->first("filePath");
      if (!fp) {
        err = "GroupObject without filePath";
        return false;
      }
      M4 gt = identity();
      if (g->first("transform") && !read_matrix(g->first("transform")->value(), gt)) {
        err = "bad GroupObject transform";
        return false;
      }
      Shape whole;
      std::vector<Shape> parts;
      if (!load_shape(resolve(fp->value()), true, true, whole, parts, err)) return false;
      for (auto& o : g->kids) {
        if (o->tag != "Object") continue;
        const XNode* mat = material(o.get(), scene);
        const char* nm = o->attr("Name");
        const Shape* s = nullptr;
        for (auto& p : parts)
          if (nm && p.name == nm) s = &p;
        if (!mat || !s) {
          err = std::string("GroupObject object ") + (nm ? nm : "?") + " has no material or group";
          return false;
        }
        rx_object_desc d;
        if (!material_desc(mat, d)) return false;
        add_object(*s, gt, d);  // the copied shape carries the group's transform
      }
    }
    for (auto& o : scene->kids) {
      if (o->tag != "Object") continue;
      const XNode* mat = material(o.get(), scene);
      const XNode* fp = o->first("filePath");
      if (!mat || !fp) {
        err = "Object without material or filePath";
        return false;
      }
      rx_object_desc d;
      if (!material_desc(mat, d)) return false;
      M4 t = identity();
      if (o->first("transform") && !read_matrix(o->first("transform")->value(), t)) {
        err = "bad Object transform";
        return false;
      }
      Shape s;
      std::vector<Shape> none;
      if (!load_shape(resolve(fp->value()), false, false, s, none, err)) return false;
      add_object(s, t, d);
    }
    const XNode* camname = sc_cfg->first("camera");
    const XNode* c = find(scene, "Camera", camname ? camname->value() : "");
    if (!c || !c->first("width") || !c->first("height") || !c->first("sightDist") || !c->first("focus") ||
        !c->first("position") || !c->first("up")) {
      err = "camera incomplete or missing";
      return false;
    }
    cam.width = atoi(c->first("width")->value());
    cam.height = atoi(c->first("height")->value());
    cam.sight_dist = (uint32_t)atoi(c->first("sightDist")->value());
    if (!read_vec(c->first("focus")->value(), cam.focus) || !read_vec(c->first("position")->value(), cam.position) ||
        !read_vec(c->first("up")->value(), cam.up)) {
      err = "bad camera vector";
      return false;
    }
    prm.spp = 1;
    prm.max_depth = 20;  // MCRenderer's default (include/MCRenderer.h:131-141)
    prm.mode = RX_MODE_PT;
    prm.seed = 1;
    std::string type;
    if (cfg->first("renderer")) type = cfg->first("renderer")->value();
    if (cfg->first("maxDepth")) prm.max_depth = (uint32_t)atoi(cfg->first("maxDepth")->value());
    if (const XNode* r = cfg->first("Renderer")) {
      if (r->first("type")) type = r->first("type")->value();
      if (r->first("maxDepth")) prm.max_depth = (uint32_t)atoi(r->first("maxDepth")->value());
      if (r->first("spp")) prm.spp = (uint32_t)atoi(r->first("spp")->value());
    }
    if (type == "PTtest") prm.mode = RX_MODE_NEE;
    else if (type == "BPT" || type == "BidirectionalPathTracer") prm.mode = RX_MODE_BPT;
    else if (!type.empty() && type != "PathTracer" && type != "PT") {
      err = "renderer " + type + " is not on this path";
      return false;
    }
    if (const XNode* tm = cfg->first("time")) {
      if (tm->first("hours")) runtime += 3600.0 * atoi(tm->first("hours")->value());
      if (tm->first("minutes")) runtime += 60.0 * atoi(tm->first("minutes")->value());
    }
    if (cfg->first("savePath")) save_path = cfg->first("savePath")->value();
    return true;
  }
};

}  // namespace

struct rx_scene_file {
  Loader L;
};

extern "C" {

int rx_scene_load(const char* config_path, rx_scene_file** out, char* err, uint32_t err_cap) {
  if (!config_path || !out) return RX_ERR_INVALID;
  *out = nullptr;
  rx_scene_file* f = new (std::nothrow) rx_scene_file;
  if (!f) return RX_ERR_NO_MEMORY;
  if (!f->L.load(config_path)) {
    if (err && err_cap) snprintf(err, err_cap, "%s", f->L.err.c_str());
    delete f;
    return RX_ERR_INVALID;
  }
  *out = f;
  return RX_OK;
}

void rx_scene_file_free(rx_scene_file* f) { delete f; }

int rx_scene_file_sizes(const rx_scene_file* f, uint32_t* n_verts, uint32_t* n_tris, uint32_t* n_objects) {
  if (!f) return RX_ERR_INVALID;
  if (n_verts) *n_verts = (uint32_t)(f->L.verts.size() / 3);
  if (n_tris) *n_tris = (uint32_t)(f->L.tris.size() / 3);
  if (n_objects) *n_objects = (uint32_t)f->L.objs.size();
  return RX_OK;
}

int rx_scene_file_get(const rx_scene_file* f, float* verts, uint32_t* tris, rx_object_desc* objects,
                      rx_camera_desc* camera, rx_render_params* params, char* save_path, uint32_t save_cap,
                      double* runtime_s) {
  if (!f) return RX_ERR_INVALID;
  const Loader& L = f->L;
  if (verts) memcpy(verts, L.verts.data(), L.verts.size() * 4);
  if (tris) memcpy(tris, L.tris.data(), L.tris.size() * 4);
  if (objects) memcpy(objects, L.objs.data(), L.objs.size() * sizeof(rx_object_desc));
  if (camera) *camera = L.cam;
  if (params) *params = L.prm;
  if (save_path && save_cap) snprintf(save_path, save_cap, "%s", L.save_path.c_str());
  if (runtime_s) *runtime_s = L.runtime;
  return RX_OK;
}

}  // extern "C"
