// mdwt_offline — offline renderer with the reference's mdwtOffline flags and frame loop
// (viewer.cpp:70-161 argument set, 486-528 Run loop, 791-828 summary + snapshot), on
// libmdwt through the C++ Renderer API (include/mdwt_renderer.hpp).
//
//   mdwt_offline -fr a_256x256x256_uint8.raw b_256x256x256_uint8.raw -t a.tf b.tf
//                -m 2 -mc 16 -r 1024 1024 -n 256 -wu 8 -o out
//
// Transfer functions (-t): the reference reads the imgui transfer-function widget's state
// files, whose format belongs to an absent submodule. Here a .tf file is text:
//   optional lines "opacity <scale>" and "range <lo> <hi>", then one "r g b a" line per
//   texel (alpha in [0,1]; it is cubed on upload as in SetXFColormap).
// Without -t every channel gets a white ramp with alpha = value.
// Lighting (-l shadows gradient dx dy dz [intensity [ambient]]) as viewer.cpp:271-281.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mdwt_renderer.hpp"

namespace {

// ---- minimal PNG writer: RGBA8, zlib stored (uncompressed) blocks ------------------------
uint32_t crc_table[256];
void crc_init() {
  for (uint32_t n = 0; n < 256; ++n) {
    uint32_t c = n;
    for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
    crc_table[n] = c;
  }
}
uint32_t crc(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
  for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xff] ^ (c >> 8);
  return c;
}
void put32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back(uint8_t(x >> 24));
  v.push_back(uint8_t(x >> 16));
  v.push_back(uint8_t(x >> 8));
  v.push_back(uint8_t(x));
}
void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
  put32(out, uint32_t(data.size()));
  std::vector<uint8_t> td(type, type + 4);
  td.insert(td.end(), data.begin(), data.end());
  out.insert(out.end(), td.begin(), td.end());
  put32(out, crc(td.data(), td.size()) ^ 0xFFFFFFFFu);
}
// rgba: row-major, row 0 = first image row of the frame buffer; written bottom-up like the
// reference (stbi_flip_vertically_on_write(true), viewer.cpp:65)
bool write_png(const std::string& path, const std::vector<uint32_t>& rgba, int w, int h) {
  crc_init();
  std::vector<uint8_t> raw;
  raw.reserve(size_t(h) * (size_t(w) * 4 + 1));
  for (int y = h - 1; y >= 0; --y) {
    raw.push_back(0);  // filter: none
    for (int x = 0; x < w; ++x) {
      const uint32_t p = rgba[size_t(y) * w + x];
      raw.push_back(uint8_t(p));
      raw.push_back(uint8_t(p >> 8));
      raw.push_back(uint8_t(p >> 16));
      raw.push_back(uint8_t(p >> 24));
    }
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  uint32_t a = 1, b = 0;
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
    const size_t len = std::min<size_t>(65535, raw.size() - off);
    z.push_back(off + len >= raw.size() ? 1 : 0);
    z.push_back(uint8_t(len));
    z.push_back(uint8_t(len >> 8));
    z.push_back(uint8_t(~len));
    z.push_back(uint8_t(~len >> 8));
    z.insert(z.end(), raw.begin() + long(off), raw.begin() + long(off + len));
    if (off + len >= raw.size()) break;
  }
  put32(z, (b << 16) | a);
  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  std::vector<uint8_t> ihdr;
  put32(ihdr, uint32_t(w));
  put32(ihdr, uint32_t(h));
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
  chunk(png, "IHDR", ihdr);
  chunk(png, "IDAT", z);
  chunk(png, "IEND", {});
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(png.data()), std::streamsize(png.size()));
  return bool(f);
}

struct Tf {
  std::vector<std::array<float, 4>> cm;
  float opacity = 1.f;
  std::array<float, 2> range{0.f, 1.f};
};

Tf load_tf(const std::string& path) {
  Tf t;
  std::ifstream f(path);
  if (!f) throw mdwt::Error(MDWT_ERR_ARG, "cannot open transfer function " + path);
  std::string line;
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    std::string w;
    if (!(ss >> w) || w[0] == '#') continue;
    if (w == "opacity") {
      ss >> t.opacity;
    } else if (w == "range") {
      ss >> t.range[0] >> t.range[1];
    } else {
      std::array<float, 4> c{};
      c[0] = std::stof(w);
      ss >> c[1] >> c[2] >> c[3];
      t.cm.push_back(c);
    }
  }
  if (t.cm.empty()) throw mdwt::Error(MDWT_ERR_ARG, path + ": no texels");
  return t;
}

Tf default_tf() {
  Tf t;
  for (int i = 0; i < 256; ++i) t.cm.push_back({1.f, 1.f, 1.f, (i + 0.5f) / 256.f});
  return t;
}

const char* mode_string(int m) {  // viewer.cpp:238-270
  switch (m) {
    case 0: return "BLN";
    case 1: return "MMB";
    case 2: return "AMMB";
    case 3: return "CMB";
    case 4: return "MAX";
    case 5: return "MIX";
    case 6: return "MM_RM";
    case 7: return "MAX_RM";
    case 8: return "MIX_RM";
    case 9: return "COMP";
    default: return "?";
  }
}

int usage(const char* msg) {
  std::fprintf(stderr,
               "%s\nusage: mdwt_offline -fr FILE... [-t TF...] [-m 0..9] [-mc M | -mc X Y Z]\n"
               "  [-r W H] [-c px py pz ax ay az ux uy uz fovy] [-bg r g b] [-dt dt] [-cb [sx sy sz]]\n"
               "  [-spp N] [-acc 0|1] [-l shadows gradient dx dy dz [intensity [ambient]]]\n"
               "  [-n frames] [-wu warmup] [-nt N] [-o name] [-seed S]\n",
               msg);
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  std::map<std::string, std::vector<std::string>> args;
  std::string key;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    const bool is_flag = a.size() > 1 && a[0] == '-' && !(std::isdigit(static_cast<unsigned char>(a[1])) || a[1] == '.');
    if (is_flag) {
      key = a;
      args[key];
    } else if (!key.empty()) {
      args[key].push_back(a);
    } else {
      return usage(("unexpected argument " + a).c_str());
    }
  }
  auto has = [&](const char* k) { return args.count(k) > 0; };
  auto f = [&](const char* k, size_t i) { return std::stof(args[k].at(i)); };
  auto u = [&](const char* k, size_t i) { return std::stoi(args[k].at(i)); };
  try {
    if (!has("-fr") || args["-fr"].empty()) return usage("No data file given");
    mdwt::Renderer r(0);
    int mode = has("-m") ? u("-m", 0) : 0;
    if (has("-r")) r.fbSize = {u("-r", 0), u("-r", 1)};
    unsigned frames = has("-n") ? unsigned(u("-n", 0)) : 100;
    unsigned wu = has("-wu") ? unsigned(u("-wu", 0)) : 0;
    unsigned nth = has("-nt") ? unsigned(u("-nt", 0)) : 0;
    std::string out = has("-o") ? args["-o"].at(0) : "";
    if (has("-c")) {
      if (args["-c"].size() != 10) return usage("-c takes 10 floats");
      r.camera.setOrientation({f("-c", 0), f("-c", 1), f("-c", 2)}, {f("-c", 3), f("-c", 4), f("-c", 5)},
                              {f("-c", 6), f("-c", 7), f("-c", 8)}, f("-c", 9));
    }
    if (has("-mc")) {  // viewer.cpp:204-224: one value is broadcast; each must be positive
      auto& mc = args["-mc"];
      if (mc.size() != 1 && mc.size() != 3) return usage("-mc takes 1 or 3 values");
      for (int i = 0; i < 3; ++i) {
        const int m = std::stoi(mc[mc.size() == 1 ? 0 : i]);
        if (m < 1) return usage("Number of macrocells per side must be positive");
        r.macrocellDims[i] = unsigned(m);
      }
    }
    if (has("-bg")) r.bgColor = {f("-bg", 0), f("-bg", 1), f("-bg", 2)};
    if (has("-spp")) r.spp = u("-spp", 0);
    if (has("-acc")) r.enableAccumulation = u("-acc", 0) != 0;
    if (has("-l")) {
      auto& l = args["-l"];
      if (l.size() < 5 || l.size() > 7) return usage("-l takes 5 to 7 values");
      r.enableShadows = std::stof(l[0]) != 0.f;
      r.enableGradientShading = std::stof(l[1]) != 0.f;
      r.lightDir = {std::stof(l[2]), std::stof(l[3]), std::stof(l[4])};
      if (l.size() > 5) r.lightIntensity = std::stof(l[5]);
      if (l.size() > 6) r.ambient = std::stof(l[6]);
    }
    for (const auto& path : args["-fr"]) {
      auto raw = std::make_shared<mdwt::RawR>(path);
      if (has("-cb")) {
        if (args["-cb"].size() == 3)
          raw->reshapeBounds({f("-cb", 0), f("-cb", 1), f("-cb", 2)});
        else if (args["-cb"].empty())
          raw->reshapeBounds();
        else
          return usage("Bounds must be given as 3 floats or not at all");
      }
      std::printf("found %d x %d x %d voxels\n", raw->dims[0], raw->dims[1], raw->dims[2]);
      if (!r.PushMesh(raw)) return usage("too many channels");
    }
    r.Init(unsigned(mode), !r.camera.given);
    if (has("-seed")) mdwt::check(mdwt_set_seed(r.handle(), uint32_t(std::stoul(args["-seed"].at(0)))), r.handle());
    if (has("-dt")) r.Setdt(f("-dt", 0));
    const size_t nch = args["-fr"].size();
    for (size_t i = 0; i < nch; ++i) {
      const Tf t = (has("-t") && i < args["-t"].size()) ? load_tf(args["-t"][i]) : default_tf();
      r.SetXFColormap(t.cm, i);
      r.SetXFOpacityScale(t.opacity, i);
      r.SetXFRange(t.range, i);
    }
    const std::string tag = std::string("_") + mode_string(mode) + "_w_";
    while (unsigned(r.frameID) < frames + wu) {  // viewer.cpp:510-528
      if (unsigned(r.frameID) <= wu) r.ResetAccumulation();
      r.Render();
      if (unsigned(r.frameID) > wu && nth != 0 && (r.frameID - wu) % nth == 0 && !out.empty())
        write_png(out + tag + std::to_string(r.frameID - wu) + std::string("_shadows_") + (r.enableShadows ? "on" : "off") + "_frames.png", r.FrameBuffer(),
                  r.fbSize[0], r.fbSize[1]);
    }
    std::printf("============ Scene Info ============\n");
    std::printf("Number of frames: %u\n", unsigned(r.frameID) - wu);
    std::printf("Macrocell dimesions: %u, %u, %u\n", r.macrocellDims[0], r.macrocellDims[1],
                r.macrocellDims[2]);
    std::printf("Mode: %s\n", mode_string(mode));
    std::printf("avg. fps: %.3f (%0.4f sec)\n", 1.0f / r.avgTime, r.avgTime);
    std::printf("best. fps: %.3f (%0.4f sec)\n", 1.0f / r.minTime, r.minTime);
    if (!out.empty())
      write_png(out + tag + std::to_string(r.frameID - int(wu)) + std::string("_shadows_") + (r.enableShadows ? "on" : "off") + "_frames.png", r.FrameBuffer(),
                r.fbSize[0], r.fbSize[1]);
  } catch (const mdwt::Error& e) {
    std::fprintf(stderr, "error (%d): %s\n", e.status, e.what());
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
