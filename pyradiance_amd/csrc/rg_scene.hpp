// Scene compilation: Radiance objects -> flat device tables.
#pragma once
#include <string>
#include <vector>
#include "rg_types.hpp"
#include "scene_io.hpp"

namespace rg {

struct CompiledScene {
  std::vector<Prim> prims;
  std::vector<double> geom;
  std::vector<Material> mats;
  std::vector<std::string> mat_names;        // name of each material object
  std::vector<Source> srcs;
  std::vector<float> lo, hi;                 // widened FP32 bounds per primitive
  std::vector<std::string> prim_modname;     // name of each primitive's modifier
  std::vector<std::string> src_modname;      // name of each source object's modifier
  std::vector<std::string> warnings;
  double bmin[3], bmax[3];
  double cusize = 0;
  double cuorg[3] = {0, 0, 0};               // octree cube as oconv sets it (ot/oconv.c:120-139)
  double cube_size = 0;
  std::vector<SrcCache> scache;              // shadow cache layout per source
  int32_t ncells = 0;                        // blocker cells of all sources
  std::vector<double> cast_rays;             // shadow-cache rays (origin, direction), in the reference's order
  std::vector<int32_t> cast_src;             // source of each cast ray
  int lamb = -1;                             // built-in Lambertian (rt/raytrace.c:24-28)
};

// Throws RadError on unsupported or malformed input.
CompiledScene compile_scene(const SceneDB& db);

}  // namespace rg
