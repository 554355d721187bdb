// View rays on the GPU: viewray() of common/image.c:213-301 for every pixel.
#include "rg_view.hpp"

#include <math.h>

namespace rg {

namespace {

constexpr double kPIv = 3.14159265358979323846;
constexpr double FTINY_V = 1e-6;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {   // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// viewray(): origin and direction for image position (x, y) in [0,1)^2;
// returns < -FTINY where the view has no ray
__device__ double view_ray(const ViewSpec& v, double x, double y, double o[3], double d[3]) {
  double z, dd;
  x += v.hoff - 0.5;
  y += v.voff - 0.5;
  switch (v.type) {
    case 'l':
      for (int i = 0; i < 3; ++i) {
        o[i] = v.vp[i] + v.vfore * v.vdir[i] + x * v.hvec[i] + y * v.vvec[i];
        d[i] = v.vdir[i];
      }
      return 0.0;
    case 'v':
      for (int i = 0; i < 3; ++i) d[i] = v.vdir[i] + x * v.hvec[i] + y * v.vvec[i];
      for (int i = 0; i < 3; ++i) o[i] = v.vp[i] + d[i] * v.vfore;
      dd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (dd > 0) for (int i = 0; i < 3; ++i) d[i] /= dd;
      return 0.0;
    case 'h':
      z = 1.0 - x * x * v.hn2 - y * y * v.vn2;
      if (z < 0.0) return -1.0;
      z = sqrt(z);
      for (int i = 0; i < 3; ++i) d[i] = z * v.vdir[i] + x * v.hvec[i] + y * v.vvec[i];
      for (int i = 0; i < 3; ++i) o[i] = v.vp[i] + d[i] * v.vfore;
      return 0.0;
    case 'c':
      dd = x * v.horiz * (kPIv / 180.0);
      z = cos(dd);
      x = sin(dd);
      for (int i = 0; i < 3; ++i) d[i] = z * v.vdir[i] + x * v.hvec[i] + y * v.vvec[i];
      for (int i = 0; i < 3; ++i) o[i] = v.vp[i] + d[i] * v.vfore;
      dd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (dd > 0) for (int i = 0; i < 3; ++i) d[i] /= dd;
      return 0.0;
    case 'a':
      x *= (1.0 / 180.0) * v.horiz;
      y *= (1.0 / 180.0) * v.vert;
      dd = x * x + y * y;
      if (dd > 1.0) return -1.0;
      dd = sqrt(dd);
      z = cos(kPIv * dd);
      dd = dd <= FTINY_V ? kPIv : sqrt(1.0 - z * z) / dd;
      x *= dd;
      y *= dd;
      for (int i = 0; i < 3; ++i) d[i] = z * v.vdir[i] + x * v.hvec[i] + y * v.vvec[i];
      for (int i = 0; i < 3; ++i) o[i] = v.vp[i] + d[i] * v.vfore;
      return 0.0;
    case 's':
      x *= sqrt(v.hn2);
      y *= sqrt(v.vn2);
      dd = x * x + y * y;
      z = (1. - dd) / (1. + dd);
      x *= (1. + z);
      y *= (1. + z);
      for (int i = 0; i < 3; ++i) d[i] = z * v.vdir[i] + x * v.hvec[i] + y * v.vvec[i];
      for (int i = 0; i < 3; ++i) o[i] = v.vp[i] + d[i] * v.vfore;
      return 0.0;
  }
  return -1.0;
}

__global__ void k_view_rays(ViewSpec v, int xres, int yres, double pj, uint64_t seed, double* __restrict__ rays,
                            uint8_t* __restrict__ valid) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)xres * yres;
  if (i >= n) return;
  const int row = (int)(i / xres), x = (int)(i % xres);
  const int y = yres - 1 - row;
  double jx = .5, jy = .5;
  if (pj > 0) {
    uint64_t h = mix64(seed * 0x9e3779b97f4a7c15ULL + (uint64_t)i + 1);
    double u1 = (double)(h >> 11) * (1.0 / 9007199254740992.0);
    double u2 = (double)(mix64(h) >> 11) * (1.0 / 9007199254740992.0);
    jx = .5 + pj * (.5 - u1);
    jy = .5 + pj * (.5 - u2);
  }
  double o[3], d[3];
  double r = view_ray(v, (x + jx) / xres, (y + jy) / yres, o, d);
  double* out = rays + 6 * i;
  bool ok = r >= -FTINY_V && (d[0] != 0 || d[1] != 0 || d[2] != 0);
  for (int k = 0; k < 3; ++k) {
    out[k] = ok ? o[k] : 0.0;
    out[3 + k] = ok ? d[k] : 0.0;
  }
  if (!ok) out[5] = 1.0;           // any direction: the pixel is cleared afterwards
  valid[i] = ok ? 1 : 0;
}

}  // namespace

void view_rays(const ViewSpec& v, int xres, int yres, double pj, uint64_t seed, double* rays, uint8_t* valid,
               hipStream_t st) {
  const long long n = (long long)xres * yres;
  const int block = 256;
  hipLaunchKernelGGL(k_view_rays, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, st, v, xres, yres, pj,
                     seed, rays, valid);
}

}  // namespace rg
