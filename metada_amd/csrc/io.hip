// File formats around the analysis (host code): the reference's observation
// text format (GridObservation::loadFromFile, GridObservation.hpp:401-476),
// its state text format (SimpleState::saveToFile / readFromFile,
// SimpleState.hpp:162-188, 248-263), and an FP64-exact binary sidecar for
// states, which the 6-decimal text cannot hold.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mda_internal.hpp"

struct mda_obs_set {
  std::vector<double> value, variance, a, b, c;
  int32_t coord = MDA_COORD_GRID;
};

namespace {

// Context-free calls (ctx == NULL) keep their last error per host thread;
// mda_last_error(NULL) returns it.
thread_local std::string g_io_error;

template <class F>
int io_guarded(mda_ctx* ctx, F&& f) {
  if (ctx) return mda::guarded_call(ctx, f);
  try {
    f();
    g_io_error.clear();
    return MDA_OK;
  } catch (const mda_error& e) {
    g_io_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_io_error = e.what();
    return MDA_E_DEVICE;
  }
}

// Binary sidecar: 64-byte little-endian header, then n doubles.
//   bytes 0-7   magic "MDASTATE"
//   bytes 8-15  version (1)
//   bytes 16-39 nx, ny, nlev (int64)
//   bytes 40-63 reserved (zero)
constexpr char SIDECAR_MAGIC[9] = "MDASTATE";
constexpr int64_t SIDECAR_VERSION = 1;

}  // namespace

extern "C" {

int mda_obs_read_grid_file(mda_ctx* ctx, const char* path, double error, double missing_value,
                           int32_t coord, mda_obs_set** out) {
  return io_guarded(ctx, [&] {
    MDA_REQUIRE(path && out, "path and out are required");
    MDA_REQUIRE(coord == MDA_COORD_GRID || coord == MDA_COORD_GEOGRAPHIC,
                "Unsupported coordinate system (grid or geographic)");
    *out = nullptr;
    std::ifstream file(path);
    if (!file.is_open()) throw mda_error(MDA_E_ARG, std::string("Could not open observation file: ") + path);
    auto set = std::make_unique<mda_obs_set>();
    set->coord = coord;
    std::string line;
    bool header = false, started = false;
    while (std::getline(file, line)) {
      if (line.empty()) continue;
      if (line.find("Time") != std::string::npos && line.find("XLONG_U") != std::string::npos) {
        header = true;
        continue;
      }
      if (line.find("---") != std::string::npos &&
          line.find_first_not_of("- \t") == std::string::npos) {
        if (header) started = true;
        continue;
      }
      if (!started) continue;
      std::istringstream iss(line);
      int time, z, y, x;
      double znu, xlong, xlat, v;
      if (!(iss >> time >> z >> y >> x >> znu >> xlong >> xlat >> v)) continue;
      if (v == missing_value) continue;
      set->value.push_back(v);
      set->variance.push_back(error * error);  // getCovariance (GridObservation.hpp:239-252)
      if (coord == MDA_COORD_GRID) {
        set->a.push_back(x); set->b.push_back(y); set->c.push_back(z);
      } else {
        set->a.push_back(xlat); set->b.push_back(xlong); set->c.push_back(znu);
      }
    }
    if (set->value.empty())
      throw mda_error(MDA_E_ARG, std::string("No valid observations loaded from file: ") + path);
    *out = set.release();
  });
}

int mda_obs_set_view(const mda_obs_set* set, mda_obs* view, int32_t* coord) {
  if (!set || !view) return MDA_E_ARG;
  std::memset(view, 0, sizeof(*view));
  view->p = (int64_t)set->value.size();
  view->value = set->value.data();
  view->variance = set->variance.data();
  view->loc_a = set->a.data();
  view->loc_b = set->b.data();
  view->loc_c = set->c.data();
  view->on_device = 0;
  if (coord) *coord = set->coord;
  return MDA_OK;
}

void mda_obs_set_free(mda_obs_set* set) { delete set; }

const char* mda_last_io_error(void) { return g_io_error.c_str(); }

int mda_state_read_text(mda_ctx* ctx, const char* path, double* out, int64_t capacity,
                        int64_t* count) {
  return io_guarded(ctx, [&] {
    MDA_REQUIRE(path && count, "path and count are required");
    std::ifstream file(path);
    if (!file) throw mda_error(MDA_E_ARG, std::string("Cannot open state file: ") + path);
    int64_t n = 0;
    std::string line;
    while (std::getline(file, line)) {
      std::istringstream iss(line);
      double v;
      while (iss >> v) {
        if (out && n < capacity) out[n] = v;
        ++n;
      }
    }
    *count = n;
    MDA_REQUIRE(!out || n <= capacity, "state file holds more values than the buffer");
  });
}

int mda_state_write_text(mda_ctx* ctx, const char* path, const double* v, int64_t nx,
                         int64_t nrows) {
  return io_guarded(ctx, [&] {
    MDA_REQUIRE(path && v && nx > 0 && nrows > 0, "path, values and dimensions are required");
    FILE* f = std::fopen(path, "w");
    if (!f) throw mda_error(MDA_E_ARG, std::string("Could not open file for writing: ") + path);
    for (int64_t y = 0; y < nrows; ++y) {
      for (int64_t x = 0; x < nx; ++x)
        std::fprintf(f, x + 1 < nx ? "%12.6f " : "%12.6f", v[y * nx + x]);
      std::fputc('\n', f);
    }
    const bool ok = std::fclose(f) == 0;
    if (!ok) throw mda_error(MDA_E_ARG, std::string("write failed: ") + path);
  });
}

int mda_state_write_binary(mda_ctx* ctx, const char* path, const double* v, int64_t nx,
                           int64_t ny, int64_t nlev) {
  return io_guarded(ctx, [&] {
    MDA_REQUIRE(path && v && nx > 0 && ny > 0 && nlev > 0, "path, values and dimensions are required");
    unsigned char hdr[64] = {0};
    std::memcpy(hdr, SIDECAR_MAGIC, 8);
    const int64_t fields[4] = {SIDECAR_VERSION, nx, ny, nlev};
    std::memcpy(hdr + 8, fields, sizeof(fields));
    FILE* f = std::fopen(path, "wb");
    if (!f) throw mda_error(MDA_E_ARG, std::string("Could not open file for writing: ") + path);
    const size_t n = (size_t)(nx * ny * nlev);
    const bool ok = std::fwrite(hdr, 1, 64, f) == 64 && std::fwrite(v, sizeof(double), n, f) == n;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) throw mda_error(MDA_E_ARG, std::string("write failed: ") + path);
  });
}

int mda_state_read_binary(mda_ctx* ctx, const char* path, double* out, int64_t capacity,
                          int64_t* nx, int64_t* ny, int64_t* nlev) {
  return io_guarded(ctx, [&] {
    MDA_REQUIRE(path && nx && ny && nlev, "path and dimension outputs are required");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw mda_error(MDA_E_ARG, std::string("Cannot open state file: ") + path);
    unsigned char hdr[64];
    const bool got = std::fread(hdr, 1, 64, f) == 64;
    int64_t fields[4] = {0, 0, 0, 0};
    if (got) std::memcpy(fields, hdr + 8, sizeof(fields));
    if (!got || std::memcmp(hdr, SIDECAR_MAGIC, 8) != 0 || fields[0] != SIDECAR_VERSION ||
        fields[1] <= 0 || fields[2] <= 0 || fields[3] <= 0) {
      std::fclose(f);
      throw mda_error(MDA_E_ARG, std::string("not a state sidecar: ") + path);
    }
    *nx = fields[1]; *ny = fields[2]; *nlev = fields[3];
    const int64_t n = fields[1] * fields[2] * fields[3];
    bool ok = true;
    if (out) {
      if (n > capacity) {
        std::fclose(f);
        throw mda_error(MDA_E_ARG, "state sidecar holds more values than the buffer");
      }
      ok = std::fread(out, sizeof(double), (size_t)n, f) == (size_t)n;
    }
    std::fclose(f);
    if (!ok) throw mda_error(MDA_E_ARG, std::string("state sidecar is short: ") + path);
  });
}

}  // extern "C"
