// Host-side Radiance scene database: .rad text and .oct octree input.
#pragma once
#include <string>
#include <vector>
#include <unordered_map>
#include <stdexcept>
#include <stdint.h>

namespace rg {

struct OctReader;

struct RadError : std::runtime_error {
  explicit RadError(const std::string& m) : std::runtime_error(m) {}
};

// One scene object as Radiance stores it (common/object.h OBJREC).
struct RObj {
  std::string type;               // type keyword ("polygon", "plastic", ...)
  int omod = -1;                  // modifier object index, -1 = void
  std::string name;
  std::vector<std::string> sargs;
  std::vector<double> fargs;
  std::string dir;                // mesh / instance objects: directory of the defining file
};

// A compiled triangle mesh (common/mesh.h MESH) with its vertices decoded:
// positions from the 32-bit cube coordinates, normals from the 4-byte
// direction code (common/mesh.c:236-278, common/dircode.c:83-106).
struct MeshData {
  std::string path;
  std::vector<double> xyz;        // [vertex][3]
  std::vector<double> nrm;        // [vertex][3], zero where the vertex has none
  std::vector<uint8_t> has_n;     // vertex carries a normal
  struct Tri { int32_t v[3]; int32_t mat; };   // flat vertex indices; material object or -1
  std::vector<Tri> tris;
  int mat0 = 0, nmats = 0;        // the mesh's own material objects in SceneDB::objs
};

class SceneDB {
 public:
  std::vector<RObj> objs;
  double cuorg[3] = {0, 0, 0};
  double cusize = 0;
  bool allow_exec = false;        // run "!command" lines (readobj.c:58-61)
  std::vector<MeshData> meshes;
  std::unordered_map<int, int> obj_mesh;    // mesh object -> meshes[]
  // Octrees placed by instance objects: their objects are appended to objs
  // (common/instance.c:23-60) as one block per file and are not part of the
  // top-level scene (hidden), so they are neither primitives nor sources
  // except through their instances.
  struct Block { int first, count; };
  std::vector<Block> inst_blocks;
  std::unordered_map<int, int> obj_inst;    // instance object -> inst_blocks[]
  std::vector<uint8_t> hidden;              // per object: inside an instanced octree

  // Loads the octree of every instance object not loaded yet (nested ones
  // too); files are shared.
  void resolve_instances();

  // Loads the mesh file of every mesh object not loaded yet, appending each
  // file's material objects after the scene's (as the renderer's preload
  // does, rt/preload.c with common/mesh.c:74-139).  Files are shared.
  void resolve_meshes();

  void load_rad_text(const std::string& text, const std::string& where);
  void load_rad_file(const std::string& path);
  void load_oct(const std::string& path);
  void load_any(const std::string& path);   // by magic: octree or scene text

  int modifier(const std::string& name) const;   // last definition, -1 if none
  static bool is_surface(const std::string& t);
  static bool is_material(const std::string& t);
  static bool is_modifier(const std::string& t) { return !is_surface(t) && t != "instance" && t != "mesh"; }

 private:
  std::unordered_map<std::string, int> modtab_;
  std::unordered_map<std::string, int> mesh_files_;
  std::unordered_map<std::string, int> inst_files_;
  std::string find_file(const RObj& o, const char* what) const;
  std::string cur_dir_;           // directory of the file being read
  void add(RObj&& o);
  void parse_stream(const std::string& text, const std::string& where, int depth);
  int load_mesh(const std::string& path);
  void read_objects(OctReader& rd, int objsize, const std::string& dir);
};

}  // namespace rg
