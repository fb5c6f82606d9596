// mdwt_renderer.hpp — C++ host API mirroring dtracker::Renderer (renderer.h:52-222, RAW
// path) on top of the C ABI in mdwt.h. Header-only; link against libmdwt.so.
//
// Public fields set before Init() (macrocellDims, fbSize, spp, enableAccumulation, bgColor,
// camera) and the method names, argument meaning and error behaviour follow the reference:
// SetXF* silently ignore an unknown channel id (renderer.cpp:1096-1097), PushMesh returns
// false past MAX_CHANNELS (renderer.cpp:1031-1044; the reference's off-by-one 13th channel
// is rejected here). Unlike the reference nothing exits or throws from inside: failures of
// the library are reported as mdwt::Error exceptions carrying the ABI status.
#pragma once
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdwt.h"

namespace mdwt {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(int status, const mdwt_renderer* h) {
  if (status != MDWT_OK) throw Error(status, mdwt_last_error(h));
}

// raw::RawR (rawFile.h:93-165): dims and dtype parsed from
// <name>_<X>x<Y>x<Z>_<dtype>.raw (rawFile.cpp:36-83); voxels kept as stored.
struct RawR {
  std::array<int, 3> dims{0, 0, 0};
  int dtype = MDWT_U8;        // unspecified -> uint8 (rawFile.h:165)
  std::vector<uint8_t> data;  // stored voxels (int8/16/32, uint32, double converted to float32)
  std::array<float, 3> boundsHi{1.f, 1.f, 1.f};  // rawFile.h:162

  explicit RawR(const std::string& fileName) {
    std::string base = fileName.substr(fileName.find_last_of("/\\") + 1);
    std::string kind = "uint8";
    std::stringstream ss(base);
    std::string tok;
    while (std::getline(ss, tok, '_')) {
      tok = tok.substr(0, tok.find('.'));
      int x, y, z;
      char c1, c2;
      std::stringstream ts(tok);
      if ((ts >> x >> c1 >> y >> c2 >> z) && c1 == 'x' && c2 == 'x' && ts.eof()) {
        dims = {x, y, z};
        continue;
      }
      if (tok == "int8" || tok == "int16" || tok == "int32" || tok == "uint8" || tok == "uint16" ||
          tok == "uint32" || tok == "float32" || tok == "float16" || tok == "double64")
        kind = tok;
    }
    if (dims[0] <= 0) throw Error(MDWT_ERR_ARG, fileName + ": no <X>x<Y>x<Z> in the name");
    const size_t n = size_t(dims[0]) * dims[1] * dims[2];
    const size_t bpv = kind == "uint8" || kind == "int8" ? 1
                       : (kind == "uint16" || kind == "int16" || kind == "float16") ? 2
                       : kind == "double64" ? 8 : 4;
    std::vector<uint8_t> buf(n * bpv);
    std::ifstream f(fileName, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(buf.data()), std::streamsize(buf.size())))
      throw Error(MDWT_ERR_ARG, fileName + ": short read");
    if (kind == "uint8" || kind == "uint16" || kind == "float16" || kind == "float32") {
      dtype = kind == "uint8" ? MDWT_U8 : kind == "uint16" ? MDWT_U16
              : kind == "float16" ? MDWT_F16 : MDWT_F32;
      data = std::move(buf);
      return;
    }
    // rawFile.cpp:100-123 conversions of the remaining types, pushed as float32
    std::vector<float> v(n);
    for (size_t i = 0; i < n; ++i) {
      const uint8_t* p = &buf[i * bpv];
      if (kind == "int8") v[i] = float(*reinterpret_cast<const int8_t*>(p)) * 255 - 128;
      else if (kind == "int16") v[i] = float(*reinterpret_cast<const int16_t*>(p)) * 65535 - 32768;
      else if (kind == "int32") v[i] = float(*reinterpret_cast<const int32_t*>(p)) * 4294967295.f - 2147483648.f;
      else if (kind == "uint32") v[i] = float(*reinterpret_cast<const uint32_t*>(p)) * 4294967295.f;
      else v[i] = float(*reinterpret_cast<const double*>(p));
    }
    dtype = MDWT_F32;
    data.resize(n * 4);
    std::memcpy(data.data(), v.data(), n * 4);
  }
  // rawFile.cpp:217-237
  void reshapeBounds() { boundsHi = {float(dims[0]), float(dims[1]), float(dims[2])}; }
  void reshapeBounds(const std::array<float, 3>& remap) { boundsHi = remap; }
};

// camera::Camera set through setOrientation (camera.h:83-110)
struct Camera {
  std::array<float, 3> from{0.f, -1.f, 0.f}, at{0.f, -1.f, -1.f}, up{0.f, 1.f, 0.f};
  float fovy = 60.f;
  bool given = false;
  void setOrientation(const std::array<float, 3>& origin, const std::array<float, 3>& interest,
                      const std::array<float, 3>& upv, float fovyInDegrees) {
    from = origin;
    at = interest;
    up = upv;
    fovy = fovyInDegrees;
    given = true;
  }
};

class Renderer {
 public:
  explicit Renderer(int hipDevice = 0) { check(mdwt_create(&h_, hipDevice), nullptr); }
  ~Renderer() { Terminate(); }
  Renderer(const Renderer&) = delete;
  Renderer& operator=(const Renderer&) = delete;

  // public fields of dtracker::Renderer (renderer.h)
  Camera camera;
  std::array<unsigned, 3> macrocellDims{0, 0, 0};
  std::array<int, 2> fbSize{1024, 1024};
  int spp = 1;
  bool enableAccumulation = true;
  std::array<float, 3> bgColor{0.f, 0.f, 0.f};
  float lightIntensity = 1.0f, ambient = 0.1f;
  bool enableShadows = false, enableGradientShading = false;   // renderer.h:121-122
  std::array<float, 3> lightDir{0.0001f, -1.f, 0.0001f};       // renderer.h:124
  // renderer.h:134-141, one per pushed channel
  struct MaterialData {
    float diffuse = 0.8f, specular = 0.3f, shininess = 8.0f, gradientStep = 0.01f;
    std::array<float, 3> specularColor{1.f, 1.f, 1.f};
  };
  std::vector<MaterialData> materials;
  float dt = 0.05f;
  int accumID = 0, frameID = 0;
  float avgTime = 0.f, minTime = std::numeric_limits<float>::max(), totalTime = 0.f;

  bool PushMesh(std::shared_ptr<RawR> raw) {
    if (raws_.size() >= MDWT_MAX_CHANNELS) return false;
    raws_.push_back(std::move(raw));
    materials.push_back(MaterialData());
    return true;
  }

  void Init(unsigned mode = 0, bool autoSetCamera = true) {
    if (raws_.empty()) throw Error(MDWT_ERR_STATE, "No data source specified!");
    for (auto& r : raws_)
      check(mdwt_push_channel(h_, r->data.data(), r->dtype, r->dims[0], r->dims[1], r->dims[2]), h_);
    const auto& hi = raws_.back()->boundsHi;
    if (hi[0] != 1.f || hi[1] != 1.f || hi[2] != 1.f) check(mdwt_set_bounds(h_, hi.data()), h_);
    const int mc[3] = {int(macrocellDims[0]), int(macrocellDims[1]), int(macrocellDims[2])};
    check(mdwt_init(h_, int(mode), mc, (autoSetCamera && !camera.given) ? 1 : 0), h_);
    int info[10];
    check(mdwt_get_info(h_, info), h_);
    macrocellDims = {unsigned(info[0]), unsigned(info[1]), unsigned(info[2])};
    Update();
    check(mdwt_set_lighting(h_, lightIntensity, ambient), h_);
    for (size_t c = 0; c < materials.size(); ++c) PushMaterial(c);
    if (camera.given) UpdateCamera();
    check(mdwt_get_dt(h_, &dt), h_);
    inited_ = true;
    ResetAccumulation();
  }

  void SetXFColormap(const std::vector<std::array<float, 4>>& cm, size_t tfID = 0) {
    if (tfID >= raws_.size()) return;
    check(mdwt_set_xf_colormap(h_, int(tfID), cm.front().data(), int(cm.size())), h_);
    ResetAccumulation();
  }
  void SetXFOpacityScale(float opacityScale, size_t tfID = 0) {
    if (tfID >= raws_.size()) return;
    check(mdwt_set_xf_opacity_scale(h_, int(tfID), opacityScale), h_);
    ResetAccumulation();
  }
  void SetXFRange(const std::array<float, 2>& xfDomain, size_t tfID = 0) {
    if (tfID >= raws_.size()) return;
    check(mdwt_set_xf_range(h_, int(tfID), xfDomain[0], xfDomain[1]), h_);
    ResetAccumulation();
  }
  void UpdateCamera() {
    check(mdwt_set_camera(h_, camera.from.data(), camera.at.data(), camera.up.data(), camera.fovy), h_);
    ResetAccumulation();
  }
  void Setdt(float newDt) {
    check(mdwt_set_dt(h_, newDt), h_);
    check(mdwt_get_dt(h_, &dt), h_);
    ResetAccumulation();
  }
  void Resetdt() {
    check(mdwt_reset_dt(h_), h_);
    check(mdwt_get_dt(h_, &dt), h_);
    ResetAccumulation();
  }
  void SetLightIntensity(float v) {
    lightIntensity = v;
    check(mdwt_set_lighting(h_, lightIntensity, ambient), h_);
    ResetAccumulation();
  }
  void SetAmbient(float v) {
    ambient = v;
    check(mdwt_set_lighting(h_, lightIntensity, ambient), h_);
    ResetAccumulation();
  }
  // renderer.cpp:1468-1474
  void SetLightDirection(const std::array<float, 3>& d) {
    lightDir = d;
    PushLight();
    ResetAccumulation();
  }
  // renderer.cpp:1195-1248 (unknown channel ids are ignored, as there)
  void SetMaterialDiffuse(float v, size_t ch = 0) { SetMaterial(ch, [&](MaterialData& m) { m.diffuse = v; }); }
  void SetMaterialSpecular(float v, size_t ch = 0) { SetMaterial(ch, [&](MaterialData& m) { m.specular = v; }); }
  void SetMaterialShininess(float v, size_t ch = 0) { SetMaterial(ch, [&](MaterialData& m) { m.shininess = v; }); }
  void SetMaterialGradientStep(float v, size_t ch = 0) {
    SetMaterial(ch, [&](MaterialData& m) { m.gradientStep = v; });
  }
  void SetMaterialSpecularColor(const std::array<float, 3>& v, size_t ch = 0) {
    SetMaterial(ch, [&](MaterialData& m) { m.specularColor = v; });
  }
  // renderer.cpp:1250-1301 (table in the library: mdwt_material_preset)
  void ApplyMaterialPreset(const std::string& name, size_t ch = 0) {
    float v[7];
    check(mdwt_material_preset(name.c_str(), v), h_);
    SetMaterial(ch, [&](MaterialData& m) {
      m.diffuse = v[0];
      m.specular = v[1];
      m.shininess = v[2];
      m.gradientStep = v[3];
      m.specularColor = {v[4], v[5], v[6]};
    });
  }
  void Resize(const std::array<int, 2>& newSize) {
    fbSize = newSize;
    Update();
  }
  // renderer.cpp:975-1008: spp, accumulation, background, light switches
  void Update() {
    check(mdwt_set_frame(h_, fbSize[0], fbSize[1], spp, enableAccumulation ? 1 : 0, bgColor.data()), h_);
    if (lightDir != pushedDir_ || enableShadows != pushedShadows_ ||
        enableGradientShading != pushedGradient_)
      PushLight();
  }
  // renderer.cpp:944-973
  void Render(short /*heatMapMode*/ = 0) {
    const auto start = std::chrono::system_clock::now();
    check(mdwt_render(h_, 1, nullptr), h_);
    const std::chrono::duration<double> el = std::chrono::system_clock::now() - start;
    totalTime += float(el.count());
    avgTime = totalTime / float(accumID + 1);
    minTime = std::min(minTime, float(el.count()));
    ++accumID;
    ++frameID;
  }
  void ResetAccumulation() {
    if (inited_) check(mdwt_reset_accumulation(h_), h_);
    accumID = 0;
    totalTime = 0.f;
    minTime = std::numeric_limits<float>::max();
    avgTime = 0.f;
  }
  void Terminate() {
    if (h_) mdwt_destroy(h_);
    h_ = nullptr;
  }
  // owlBufferGetPointer(frameBuffer): RGBA8, row-major, (0,0) = first pixel
  std::vector<uint32_t> FrameBuffer() {
    std::vector<uint32_t> fb(size_t(fbSize[0]) * fbSize[1]);
    check(mdwt_read_fb(h_, fb.data()), h_);
    return fb;
  }
  std::vector<float> AccumBuffer() {
    std::vector<float> a(size_t(fbSize[0]) * fbSize[1] * 4);
    check(mdwt_read_accum(h_, a.data(), nullptr), h_);
    return a;
  }
  mdwt_renderer* handle() { return h_; }

 private:
  void PushLight() {
    check(mdwt_set_light(h_, lightDir.data(), enableShadows ? 1 : 0, enableGradientShading ? 1 : 0), h_);
    pushedDir_ = lightDir;
    pushedShadows_ = enableShadows;
    pushedGradient_ = enableGradientShading;
  }
  void PushMaterial(size_t c) {
    const MaterialData& m = materials[c];
    check(mdwt_set_material(h_, int(c), m.diffuse, m.specular, m.shininess, m.gradientStep,
                            m.specularColor.data()),
          h_);
  }
  template <class F>
  void SetMaterial(size_t ch, F&& f) {
    if (ch >= materials.size()) return;
    f(materials[ch]);
    if (inited_) PushMaterial(ch);
    ResetAccumulation();
  }
  std::array<float, 3> pushedDir_{0.0001f, -1.f, 0.0001f};
  bool pushedShadows_ = false, pushedGradient_ = false;
  mdwt_renderer* h_ = nullptr;
  std::vector<std::shared_ptr<RawR>> raws_;
  bool inited_ = false;
};

}  // namespace mdwt
