// viewer.cpp — the multiserver viewer's command front end on the served form.
//
// ViewerClientServer::handle (multiserver/ViewerClientServer.cpp:63-252) takes
// command lines (permute_pixels, max_rays_per_packet, commit, json, window,
// render, renderMany) and answers with fixed replies and error texts.  Here
// the same lines drive the library: a json command's Datasets resolve names
// (or file names) against device-resident volumes the caller registered, its
// Visualization becomes VolumeVis structs with resampled TFs and its Lighting
// (Lighting.cpp:58-115), its Camera (Camera.cpp LoadFromJSON) and Renderer
// epsilon are kept; render submits one render request to a live server
// (gxy_server_*) that holds the visualization resident, and waits for it;
// renderMany N submits N requests at once, which the server traces together.
// The server is rebuilt when the visualization or the lighting resolved for
// the camera changes (Rendering::resolve_lights, Rendering.cpp:157-216).
// Only host code: every device operation goes through the C-ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <chrono>
#include <vector>

#include "../../include/gxy_hip.h"

#define GXY_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

// ------------------------------------------------------------------ JSON
struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal *get(const char *k) const {
    if (kind != OBJ) return nullptr;
    for (auto &kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct JParser {
  const char *p, *e;
  int depth = 0;
  void ws() {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
  }
  bool lit(const char *s) {
    const size_t n = strlen(s);
    if ((size_t)(e - p) < n || strncmp(p, s, n) != 0) return false;
    p += n;
    return true;
  }
  bool string(std::string &out) {
    if (p >= e || *p != '"') return false;
    p++;
    while (p < e && *p != '"') {
      if (*p == '\\') {
        if (++p >= e) return false;
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u':
            if (e - p < 5) return false;
            out += '?';   // non-ASCII escapes do not occur in state files
            p += 4;
            break;
          default: out += *p;
        }
        p++;
      } else {
        out += *p++;
      }
    }
    if (p >= e) return false;
    p++;
    return true;
  }
  bool value(JVal &v) {
    if (++depth > 64) return false;
    ws();
    if (p >= e) return false;
    bool ok = true;
    if (*p == '{') {
      v.kind = JVal::OBJ;
      p++;
      ws();
      if (p < e && *p == '}') {
        p++;
      } else {
        for (;;) {
          ws();
          std::string k;
          JVal c;
          if (!string(k)) return false;
          ws();
          if (p >= e || *p != ':') return false;
          p++;
          if (!value(c)) return false;
          v.obj.emplace_back(std::move(k), std::move(c));
          ws();
          if (p < e && *p == ',') { p++; continue; }
          if (p < e && *p == '}') { p++; break; }
          return false;
        }
      }
    } else if (*p == '[') {
      v.kind = JVal::ARR;
      p++;
      ws();
      if (p < e && *p == ']') {
        p++;
      } else {
        for (;;) {
          JVal c;
          if (!value(c)) return false;
          v.arr.push_back(std::move(c));
          ws();
          if (p < e && *p == ',') { p++; continue; }
          if (p < e && *p == ']') { p++; break; }
          return false;
        }
      }
    } else if (*p == '"') {
      v.kind = JVal::STR;
      ok = string(v.str);
    } else if (lit("true")) {
      v.kind = JVal::BOOL;
      v.b = true;
    } else if (lit("false")) {
      v.kind = JVal::BOOL;
    } else if (lit("null")) {
      v.kind = JVal::NUL;
    } else {
      char *end = nullptr;
      v.num = strtod(p, &end);
      if (end == p || end > e) return false;
      v.kind = JVal::NUM;
      p = end;
    }
    depth--;
    return ok;
  }
};

bool parse_json(const std::string &s, JVal &v) {
  JParser P{s.data(), s.data() + s.size()};
  if (!P.value(v)) return false;
  P.ws();
  return P.p == P.e;
}

bool num(const JVal *v, double &out) {
  if (!v) return false;
  if (v->kind == JVal::NUM) { out = v->num; return true; }
  if (v->kind == JVal::BOOL) { out = v->b ? 1.0 : 0.0; return true; }
  return false;
}
bool vec3(const JVal *v, float out[3]) {
  if (!v || v->kind != JVal::ARR || v->arr.size() < 3) return false;
  for (int k = 0; k < 3; k++) {
    double d;
    if (!num(&v->arr[k], d)) return false;
    out[k] = (float)d;
  }
  return true;
}
// an array of rows of `width` numbers, flattened
bool rows(const JVal *v, int width, std::vector<float> &out) {
  if (!v || v->kind != JVal::ARR) return false;
  out.clear();
  for (auto &r : v->arr) {
    if (r.kind != JVal::ARR || (int)r.arr.size() < width) return false;
    for (int k = 0; k < width; k++) {
      double d;
      if (!num(&r.arr[k], d)) return false;
      out.push_back((float)d);
    }
  }
  return true;
}

// ------------------------------------------------------------------ viewer state
struct VolEntry {
  void *handle = nullptr;
  gxy_box3f local, global;
};

struct VisSpec {   // one VolumeVis operator
  std::string dataset;
  std::vector<float> cmap, omap, range, slices, iso;
  bool dvr = false;
};

struct LightSpec {   // Lighting::LoadStateFromValue (Lighting.cpp:58-115)
  std::vector<float> src{1.f, 1.f, 1.f};   // Lighting.ispc:39-43
  std::vector<int> type{2};
  float Ka = 0.5f, Kd = 0.5f, aoR = 1.f;
  int nAO = 0, shadows = 0;
};

struct CamSpec {
  float eye[3], center[3], up[3], aov = 30.f;
};

struct ViewerObj {
  std::map<std::string, VolEntry> catalog;   // name or file name -> device volume
  std::map<std::string, VolEntry> dsets;     // dataset name -> volume (after a Datasets command)
  std::vector<VisSpec> ops;
  LightSpec light;
  CamSpec cam;
  bool have_vis = false, have_cam = false;
  float eps = 0.001f;
  int width = 512, height = 512, permute = 0, maxlist = 1000000;
  long long budget = 64LL << 30;
  int slots = 64;
  // what the live server holds
  void *server = nullptr;
  bool stale = true;                         // the visualization changed since the server was built
  std::vector<void *> tfs;
  std::vector<VolumeVis_ispc> vvs;
  std::vector<VolumeVis_ispc *> vvp;
  std::vector<std::vector<float>> sl, iv;
  Visualization_ispc V;
  Lighting_ispc L;
  std::vector<float> lsrc;
  std::vector<int> ltype;
  // frames
  std::vector<float> frame;
  int frames = 0, frame_w = 0, frame_h = 0;
  ~ViewerObj() { release(); }
  void release() {
    if (server) gxy_server_destroy(server);
    server = nullptr;
    for (void *t : tfs) gxy_tf_destroy(t);
    tfs.clear();
  }
};

std::set<void *> g_viewers;

// MappedVis defaults (MappedVis.cpp:58-62)
const float kDefaultCmap[8] = {0.f, 0.4f, 0.4f, 0.4f, 1.f, 1.f, 1.f, 1.f};
const float kDefaultOmap[4] = {0.f, 1.f, 1.f, 1.f};

std::string load_lighting(const JVal &v, LightSpec &L) {
  if (const JVal *s = v.get("Sources")) {
    if (s->kind != JVal::ARR) return "bad Sources";
    L.src.clear();
    L.type.clear();
    for (auto &e : s->arr) {
      if (e.kind != JVal::ARR || e.arr.size() < 3) return "bad Sources";
      double x, y, z, t = 0;
      if (!num(&e.arr[0], x) || !num(&e.arr[1], y) || !num(&e.arr[2], z)) return "bad Sources";
      if (e.arr.size() > 3 && !num(&e.arr[3], t)) return "bad Sources";
      float fx = (float)x, fy = (float)y, fz = (float)z;
      if ((int)t == 0) {   // directional: normalized at load (Lighting.cpp:80-93)
        const float d = std::sqrt((fx * fx + fy * fy) + fz * fz);
        if (d == 0.f) fx = fy = fz = 0.577350f;
        else { fx /= d; fy /= d; fz /= d; }
      }
      L.src.insert(L.src.end(), {fx, fy, fz});
      L.type.push_back((int)t);
    }
  }
  double d;
  if (num(v.get("shadows"), d)) L.shadows = d != 0.0;
  if (num(v.get("ao count"), d)) L.nAO = (int)d;
  if (num(v.get("ao radius"), d)) L.aoR = (float)d;
  if (num(v.get("Ka"), d)) L.Ka = (float)d;
  if (num(v.get("Kd"), d)) L.Kd = (float)d;
  return "";
}

// Visualization::LoadFromJSON: the VolumeVis operators and the Lighting
bool load_vis(ViewerObj *W, const JVal &v, std::vector<VisSpec> &ops, LightSpec &light) {
  if (v.kind != JVal::OBJ) return false;
  light = LightSpec();
  if (const JVal *l = v.get("Lighting"))
    if (!load_lighting(*l, light).empty()) return false;
  ops.clear();
  if (const JVal *o = v.get("operators")) {
    if (o->kind != JVal::ARR) return false;
    for (auto &op : o->arr) {
      const JVal *t = op.get("type");
      if (!t || t->kind != JVal::STR) return false;
      if (t->str != "Volume" && t->str != "VolumeVis") continue;
      const JVal *ds = op.get("dataset");
      if (!ds || ds->kind != JVal::STR || !W->dsets.count(ds->str)) return false;
      VisSpec s;
      s.dataset = ds->str;
      if (op.get("colormap") && !rows(op.get("colormap"), 4, s.cmap)) return false;
      if (op.get("opacitymap") && !rows(op.get("opacitymap"), 2, s.omap)) return false;
      if (const JVal *r = op.get("data range")) {
        double a, b;
        if (r->kind != JVal::ARR || r->arr.size() < 2 || !num(&r->arr[0], a) || !num(&r->arr[1], b)) return false;
        s.range = {(float)a, (float)b};
      }
      if (const JVal *iv = op.get("isovalues")) {
        if (iv->kind != JVal::ARR) return false;
        for (auto &x : iv->arr) {
          double d;
          if (!num(&x, d)) return false;
          s.iso.push_back((float)d);
        }
      }
      if (op.get("slices")) {
        if (!rows(op.get("slices"), 4, s.slices)) return false;
      } else if (const JVal *pl = op.get("plane")) {
        std::vector<float> one;
        JVal wrap;
        wrap.kind = JVal::ARR;
        wrap.arr.push_back(*pl);
        if (!rows(&wrap, 4, one)) return false;
        s.slices = one;
      }
      double dv;
      if (num(op.get("volume rendering"), dv)) s.dvr = dv != 0.0;
      ops.push_back(std::move(s));
    }
  }
  return true;
}

bool load_camera(const JVal &c, CamSpec &cam) {
  if (c.kind != JVal::OBJ) return false;
  if (!vec3(c.get("viewpoint"), cam.eye) || !vec3(c.get("viewup"), cam.up)) return false;
  if (c.get("viewcenter")) {
    if (!vec3(c.get("viewcenter"), cam.center)) return false;
  } else {
    float d[3];
    if (!vec3(c.get("viewdirection"), d)) return false;
    for (int k = 0; k < 3; k++) cam.center[k] = cam.eye[k] + d[k];
  }
  double a = 30.0;
  if (c.get("aov") && !num(c.get("aov"), a)) return false;
  cam.aov = (float)a;
  return true;
}

// Rendering::resolve_lights (Rendering.cpp:157-216): camera-relative sources
// (type 1) become points in world space.
void resolve_lights(const LightSpec &L, const CamSpec &c, std::vector<float> &src, std::vector<int> &typ) {
  auto norm = [](float a[3]) {
    float d = std::sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
    if (d != 0.f) {
      d = 1.0f / d;
      a[0] *= d; a[1] *= d; a[2] *= d;
    }
  };
  float vu[3] = {c.up[0], c.up[1], c.up[2]};
  norm(vu);
  float vd[3] = {c.center[0] - c.eye[0], c.center[1] - c.eye[1], c.center[2] - c.eye[2]};
  norm(vd);
  float right[3] = {vd[1] * vu[2] - vd[2] * vu[1], vd[2] * vu[0] - vd[0] * vu[2], vd[0] * vu[1] - vd[1] * vu[0]};
  norm(right);
  const float up2[3] = {vd[1] * right[2] - vd[2] * right[1], vd[2] * right[0] - vd[0] * right[2],
                        vd[0] * right[1] - vd[1] * right[0]};
  src = L.src;
  typ = L.type;
  for (size_t i = 0; i < typ.size(); i++)
    if (typ[i] == 1) {
      const float l[3] = {src[3 * i], src[3 * i + 1], src[3 * i + 2]};
      for (int k = 0; k < 3; k++) {
        float o = c.eye[k] + l[0] * right[k];
        o = o + l[1] * up2[k];
        src[3 * i + k] = o + l[2] * vd[k];
      }
      typ[i] = 2;
    }
}

// The visualization on the device and a live server holding it.
int build(ViewerObj *W) {
  W->release();
  W->vvs.assign(W->ops.size(), VolumeVis_ispc());
  W->vvp.clear();
  W->sl.assign(W->ops.size(), {});
  W->iv.assign(W->ops.size(), {});
  for (size_t k = 0; k < W->ops.size(); k++) {
    const VisSpec &s = W->ops[k];
    const VolEntry &vol = W->dsets[s.dataset];
    if (!vol.handle) return -1;
    const std::vector<float> cm = s.cmap.empty() ? std::vector<float>(kDefaultCmap, kDefaultCmap + 8) : s.cmap;
    const std::vector<float> om = s.omap.empty() ? std::vector<float>(kDefaultOmap, kDefaultOmap + 4) : s.omap;
    void *tf = gxy_tf_create_from_maps(cm.data(), (int)cm.size() / 4, om.data(), (int)om.size() / 2,
                                       s.range.empty() ? nullptr : s.range.data());
    if (!tf) return -1;
    W->tfs.push_back(tf);
    W->sl[k] = s.slices;
    W->iv[k] = s.iso;
    VolumeVis_ispc &v = W->vvs[k];
    memset(&v, 0, sizeof v);
    v.vis.vis.data = vol.handle;
    v.vis.transferFunction = tf;
    v.nSlices = (int)W->sl[k].size() / 4;
    v.slices = W->sl[k].empty() ? nullptr : W->sl[k].data();
    v.nIsovalues = (int)W->iv[k].size();
    v.isovalues = W->iv[k].empty() ? nullptr : W->iv[k].data();
    v.volume_render = s.dvr;
  }
  for (auto &v : W->vvs) W->vvp.push_back(&v);
  memset(&W->V, 0, sizeof W->V);
  W->V.model = nullptr;
  W->V.nVolumeVis = (int)W->vvp.size();
  W->V.volumeVis = W->vvp.empty() ? nullptr : W->vvp.data();
  // Visualization::local_commit takes the first operator's boxes (Visualization.cpp:150-162)
  if (!W->ops.empty()) {
    const VolEntry &v0 = W->dsets[W->ops[0].dataset];
    W->V.local_bb = v0.local;
    W->V.global_bb = v0.global;
  }
  resolve_lights(W->light, W->cam, W->lsrc, W->ltype);
  memset(&W->L, 0, sizeof W->L);
  W->L.n_ao_rays = W->light.nAO;
  W->L.do_shadows = W->light.shadows;
  W->L.ao_radius = W->light.aoR;
  W->L.Ka = W->light.Ka;
  W->L.Kd = W->light.Kd;
  W->L.types = W->ltype.data();
  W->L.lights = W->lsrc.data();
  W->L.nLights = (uint32_t)W->ltype.size();
  W->server = gxy_server_create(W->budget, W->slots, 256);
  if (!W->server) return -1;
  if (gxy_server_add_dataset(W->server, &W->V, W->eps) < 0) return -1;
  if (gxy_server_set_lighting(W->server, &W->L, W->eps) != 0) return -1;
  if (gxy_server_start(W->server) != 0) return -1;
  W->stale = false;
  return 0;
}

// n render requests of the current camera and window, submitted together;
// the last one's frame is kept
int render(ViewerObj *W, int n) {
  if (!W->have_vis || !W->have_cam || n <= 0) return 0;
  std::vector<float> src;
  std::vector<int> typ;
  resolve_lights(W->light, W->cam, src, typ);
  if (W->stale || !W->server || src != W->lsrc || typ != W->ltype)
    if (build(W) != 0) return -1;
  const size_t px = (size_t)W->width * W->height;
  std::vector<std::vector<float>> outs(n, std::vector<float>(px * 4, 0.f));
  std::vector<int> ids;
  for (int k = 0; k < n; k++) {
    gxy_request r;
    memset(&r, 0, sizeof r);
    r.kind = 0;
    r.dataset = 0;
    r.width = W->width;
    r.height = W->height;
    for (int a = 0; a < 3; a++) { r.eye[a] = W->cam.eye[a]; r.center[a] = W->cam.center[a]; r.up[a] = W->cam.up[a]; }
    r.aov = W->cam.aov;
    r.out = outs[k].data();
    const int id = gxy_server_submit(W->server, &r);
    if (id < 0) return -1;
    ids.push_back(id);
  }
  if (gxy_server_drain(W->server, 600.0) != 0) return -1;
  for (int id : ids) {
    gxy_request_result res;
    if (gxy_server_poll(W->server, id, &res) != 0) return -1;
  }
  W->frame.swap(outs.back());
  W->frame_w = W->width;
  W->frame_h = W->height;
  W->frames += n;
  return 0;
}

std::string handle(ViewerObj *W, const std::string &line, bool &handled) {
  handled = true;
  std::stringstream ss(line);
  std::string cmd;
  ss >> cmd;
  if (cmd == "permute_pixels") {
    std::string onoff;
    ss >> onoff;
    if (ss.fail() || onoff == "on") { W->permute = 1; return "ok"; }
    if (onoff == "off") { W->permute = 0; return "ok"; }
    return "error non-default arg to permute_pixels must be on or off";
  }
  if (cmd == "max_rays_per_packet") {
    int n;
    ss >> n;
    if (ss.fail()) return "max_rays_per_packet requires integer arg";
    W->maxlist = n;
    return "ok";
  }
  if (cmd == "commit") {
    W->stale = true;
    return "ok";
  }
  if (cmd == "json") {
    std::string json, tmp;
    while (std::getline(ss, tmp)) json += tmp;
    JVal doc;
    if (!parse_json(json, doc) || doc.kind != JVal::OBJ) return "error parsing json command";
    if (const JVal *d = doc.get("Datasets")) {
      if (d->kind != JVal::ARR) return "error loading datasets from json command";
      std::map<std::string, VolEntry> add;
      for (auto &e : d->arr) {
        const JVal *nm = e.get("name"), *fn = e.get("filename");
        if (!nm || nm->kind != JVal::STR) return "error loading datasets from json command";
        auto it = W->catalog.end();
        if (fn && fn->kind == JVal::STR) it = W->catalog.find(fn->str);
        if (it == W->catalog.end()) it = W->catalog.find(nm->str);
        if (it == W->catalog.end()) return "error loading datasets from json command";
        add[nm->str] = it->second;
      }
      for (auto &kv : add) W->dsets[kv.first] = kv.second;
      W->stale = true;
      return "ok";
    }
    if (const JVal *v = doc.get("Visualization")) {
      std::vector<VisSpec> ops;
      LightSpec light;
      if (!load_vis(W, *v, ops, light)) return "error loading visualization from json command";
      W->ops = std::move(ops);
      W->light = light;
      W->have_vis = true;
      W->stale = true;
    }
    if (const JVal *c = doc.get("Camera")) {
      CamSpec cam;
      if (!load_camera(*c, cam)) return "error loading camera from json command";
      W->cam = cam;
      W->have_cam = true;
    }
    if (const JVal *r = doc.get("Renderer")) {
      double e;
      if (r->kind == JVal::OBJ && num(r->get("epsilon"), e)) {
        W->eps = (float)e;
        W->stale = true;
      }
    }
    return "ok";
  }
  if (cmd == "window") {
    int w, h;
    ss >> w >> h;
    if (ss.fail()) return "error window command needs width and height as integers";
    W->width = w;
    W->height = h;
    return "ok";
  }
  if (cmd == "render") {
    if (render(W, 1) != 0) return std::string("error rendering: ") + gxy_last_error();
    return "ok";
  }
  if (cmd == "renderMany") {
    int n = 0;
    ss >> n;
    if (ss.fail()) n = 0;
    if (render(W, n) != 0) return std::string("error rendering: ") + gxy_last_error();
    return "ok";
  }
  handled = false;
  return "";
}

ViewerObj *viewer_of(void *v) { return g_viewers.count(v) ? (ViewerObj *)v : nullptr; }

}  // namespace

// No device work until a render: commands can be parsed and answered without a GPU.
GXY_EXPORT void *gxy_viewer_create(long long hbm_budget, int slots) {
  ViewerObj *W = new ViewerObj;
  if (hbm_budget > 0) W->budget = hbm_budget;
  if (slots > 0) W->slots = slots;
  g_viewers.insert(W);
  return W;
}

GXY_EXPORT int gxy_viewer_destroy(void *v) {
  ViewerObj *W = viewer_of(v);
  if (!W) return -1;
  g_viewers.erase(W);
  delete W;
  return 0;
}

GXY_EXPORT int gxy_viewer_add_volume(void *v, const char *key, void *volume, const gxy_box3f *local_box,
                                     const gxy_box3f *global_box) {
  ViewerObj *W = viewer_of(v);
  if (!W || !key || !local_box || !global_box) return -1;
  VolEntry e;
  e.handle = volume;
  e.local = *local_box;
  e.global = *global_box;
  W->catalog[key] = e;
  return 0;
}

GXY_EXPORT int gxy_viewer_handle(void *v, const char *line, char *reply, size_t capacity) {
  ViewerObj *W = viewer_of(v);
  if (!W || !line) return -1;
  bool handled = false;
  const std::string r = handle(W, line, handled);
  if (reply && capacity > 0) {
    const size_t n = std::min(capacity - 1, r.size());
    memcpy(reply, r.data(), n);
    reply[n] = 0;
  }
  return handled ? 1 : 0;
}

GXY_EXPORT int gxy_viewer_frame(void *v, float *out, int *width, int *height, int *frames) {
  ViewerObj *W = viewer_of(v);
  if (!W) return -1;
  if (width) *width = W->frame_w;
  if (height) *height = W->frame_h;
  if (frames) *frames = W->frames;
  if (out && !W->frame.empty()) memcpy(out, W->frame.data(), W->frame.size() * sizeof(float));
  return W->frame.empty() ? 0 : 1;
}

GXY_EXPORT int gxy_viewer_settings(void *v, int *permute, int *max_rays, int *width, int *height) {
  ViewerObj *W = viewer_of(v);
  if (!W) return -1;
  if (permute) *permute = W->permute;
  if (max_rays) *max_rays = W->maxlist;
  if (width) *width = W->width;
  if (height) *height = W->height;
  return 0;
}
