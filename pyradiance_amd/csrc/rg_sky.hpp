// Annual sky matrices from weather data: gendaymtx (gen/gendaymtx.c) with
// the Perez all-weather sky on the GPU.
//
// The host reads the weather file (WEA or EPW, gen/loadEPW.c), places the
// sun (gen/sun.c) and derives each time step's Perez parameters and
// illuminances (a few dozen scalar operations per step); the GPU evaluates
// every sky patch of every step, normalises each step's sky, spreads the sun
// over its nearest patches and writes the matrix as the reference orders it:
// [patch][time step][RGB] in FP32 (the reference's float matrix).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace rg {

struct SkyOpts {
  int mfactor = 1;                    // -m: Reinhart subdivision
  double sky_color[3] = {.96, 1.004, 1.118};   // -c (0 with -d / -n)
  double ground_color[3] = {.2, .2, .2};       // -g (0 with -d / -n)
  double sun_color[3] = {1, 1, 1};    // 0 with -s
  double rotation = 0;                // -r degrees
  int nsuns = 4;                      // sun spread over this many patches (1 with -5)
  double fixed_sun_sa = -1;           // -5: fixed solar solid angle
  int output = 0;                     // -O0 visible, -O1 solar radiance
  int sun_hours_only = 0;             // -u
  int average = 0;                    // -A
};

// Per stored time step, uploaded to the device.
struct SkyStep {
  double sun_zenith, azimuth, altitude;
  double perez[5];
  double diff_illum, dir_illum;
  float ground[3];
  int32_t kind;       // 0: zero row, 1: ground and sun only, 2: full Perez sky
  int32_t add_sun;    // AddDirect applies
  int32_t pad;
};

// One weather record that places a sun (for gendaymtx -D): its id
// (solar minute), whether it is a "good" sun, position and illuminance.
struct SkySun {
  int id;
  int good;
  double altitude, azimuth, dir_illum;
};

struct SkyPlan {
  int npatch = 0;
  int ncols = 0;                      // stored columns (1 with -A)
  int nsteps = 0;                     // stored time steps before averaging
  int ntsteps = 0;                    // time steps counted by the reference (averaging divisor)
  double latitude = 0, longitude = 0; // degrees north, east (LATLONG header)
  std::vector<float> palt, pazi, dom; // patch altitude, azimuth (radians), solid angle
  std::vector<SkyStep> steps;
  std::vector<SkySun> suns;
  SkyOpts opts;
};

// Parses weather text and prepares every step; throws std::runtime_error.
SkyPlan sky_plan(const std::string& weather, const SkyOpts& o);

// Writes the npatch x ncols x 3 matrix to device memory `out`.
void sky_matrix(const SkyPlan& plan, float* out, hipStream_t st);

// Size of the fixed device array the allocation-free form uses (per GPU).
constexpr size_t kSkyScratchBytes = size_t(96) << 20;

// sky_matrix() with its buffers in that array: no device allocation; throws
// when the plan needs more.  Calls on one device are serialised.
void sky_matrix_noalloc(const SkyPlan& plan, float* out, hipStream_t st);

}  // namespace rg
