// Primary rays of a view, generated on the GPU for picture rendering.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rg {

// A view after setview() (common/image.c:23-118): unit direction, scaled
// horizontal / vertical vectors and the squared half-size terms.
struct ViewSpec {
  int32_t type;                  // 'v' 'l' 'c' 'a' 'h' 's'
  int32_t pad;
  double vp[3], vdir[3], hvec[3], vvec[3];
  double hn2, vn2, horiz, vert, hoff, voff, vfore, vaft;
};

// One ray per pixel, scanlines top first (the picture order of rpict),
// through (x + 1/2 + pj (1/2 - u)) / xres, likewise in y, as rpict's
// pixvalue() (rt/rpict.c:668-712) places them.  valid[i] = 0 where the
// view has no ray (outside a fisheye).
void view_rays(const ViewSpec& v, int xres, int yres, double pj, uint64_t seed, double* rays, uint8_t* valid,
               hipStream_t st);

}  // namespace rg
