// Trace of the reference brick topology and region tracker under a seeded
// command sequence (test infrastructure: built from the reference's headers
// in place by oracle/build_ref.sh, output into oracle/_ref/).  One JSON object
// per line: {"op", "args", "ret", "err"}; err is [type, message] for a
// std::invalid_argument ("ValueError") or std::runtime_error ("RuntimeError").
#include "regions.hpp"

#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

using namespace fluxfx;

static std::string esc(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o;
}

static void emit(const char* op, const std::string& args, const std::string& ret, const char* type,
                 const std::string& msg) {
  if (type)
    std::printf("{\"op\":\"%s\",\"args\":%s,\"ret\":null,\"err\":[\"%s\",\"%s\"]}\n", op, args.c_str(), type,
                esc(msg).c_str());
  else
    std::printf("{\"op\":\"%s\",\"args\":%s,\"ret\":%s,\"err\":null}\n", op, args.c_str(), ret.c_str());
}

template <class F>
static void run(const char* op, const std::string& args, F&& f) {
  try {
    emit(op, args, f(), nullptr, "");
  } catch (const std::invalid_argument& e) {
    emit(op, args, "", "ValueError", e.what());
  } catch (const std::runtime_error& e) {
    emit(op, args, "", "RuntimeError", e.what());
  }
}

int main() {
  const uint32_t capacity = 48;
  BrickTopology t(capacity);
  RegionTracker r(capacity);
  std::mt19937 rng(20260);
  auto uni = [&](int lo, int hi) { return lo + int(rng() % uint32_t(hi - lo + 1)); };
  auto real = [&](double lo, double hi) { return lo + (hi - lo) * (double(rng()) / 4294967296.0); };
  auto coord = [&]() {
    int pick = uni(0, 39);
    if (pick == 0) return BrickCoord{1048575, uni(-2, 2), -1048576};
    if (pick == 1) return BrickCoord{uni(-2, 2), 1048576, 0};  // out of range
    return BrickCoord{uni(-3, 4), uni(-3, 4), uni(-2, 3)};
  };
  auto cs = [](BrickCoord c) {
    return "[" + std::to_string(c.x) + "," + std::to_string(c.y) + "," + std::to_string(c.z) + "]";
  };
  std::printf("{\"op\":\"create\",\"args\":[%u],\"ret\":null,\"err\":null}\n", capacity);
  for (int k = 0; k < 2500; ++k) {
    const int kind = uni(0, 99);
    if (kind < 34) {
      BrickCoord c = coord();
      run("activate", cs(c), [&] { return std::to_string(t.activate(c)); });
    } else if (kind < 58) {
      BrickCoord c = coord();
      run("deactivate", cs(c), [&] { return std::string(t.deactivate(c) ? "true" : "false"); });
    } else if (kind < 66) {
      BrickCoord c = coord();
      run("lookup", cs(c), [&] { return std::to_string(t.lookup(c)); });
    } else if (kind < 76) {
      BrickCoord c = coord();
      run("neighbors", cs(c), [&] {
        auto n = t.neighbors(c);
        std::string s = "[";
        for (int d = 0; d < 6; ++d) s += std::to_string(n[d]) + (d < 5 ? "," : "]");
        return s;
      });
    } else if (kind < 84) {
      BrickCoord lo = coord(), ext{uni(1, 3), uni(1, 3), uni(1, 3)};
      if (uni(0, 9) == 0) ext = {uni(0, 5), 4, 4};
      run("activate_box", "[" + cs(lo) + "," + cs(ext) + "]", [&] {
        t.activate_box(lo, ext);
        return std::string("null");
      });
    } else if (kind < 97) {
      const int n = uni(0, 3);
      std::vector<RegionSource> src;
      std::string a = "[[";
      for (int i = 0; i < n; ++i) {
        RegionSource s{real(0, 1), real(0, 1), real(0, 1), real(0, .15), real(-2, 2), real(-2, 2), real(-2, 2)};
        if (uni(0, 29) == 0) s.radius = -0.1;
        src.push_back(s);
        char buf[256];
        std::snprintf(buf, sizeof buf, "[%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g]", s.x, s.y, s.z, s.radius, s.vx,
                      s.vy, s.vz);
        a += std::string(i ? "," : "") + buf;
      }
      const int grid = uni(0, 19) == 0 ? 65 : uni(3, 8), linger = uni(1, 4), halo = uni(0, 2);
      const double dt = real(.05, .25);
      char buf[128];
      std::snprintf(buf, sizeof buf, "],%d,%.17g,%d,%d]", grid, dt, linger, halo);
      a += buf;
      run("update_regions", a, [&] {
        r.update(t, src, uint32_t(grid), dt, uint32_t(linger), uint32_t(halo));
        return std::string("null");
      });
    } else {
      std::string s = "[";
      bool first = true;
      for (auto slot : t.active()) {
        auto c = t.coord(slot);
        s += std::string(first ? "" : ",") + "[" + std::to_string(c.x) + "," + std::to_string(c.y) + "," +
             std::to_string(c.z) + "," + std::to_string(slot) + "," + std::to_string(r.age(slot)) + "]";
        first = false;
      }
      s += "]";
      run("snapshot", "[]", [&] { return s; });
    }
  }
  return 0;
}
