// Host-side Radiance scene database: .rad text and .oct octree input.
#pragma once
#include <string>
#include <vector>
#include <unordered_map>
#include <stdexcept>

namespace rg {

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
};

class SceneDB {
 public:
  std::vector<RObj> objs;
  double cuorg[3] = {0, 0, 0};
  double cusize = 0;
  bool allow_exec = false;        // run "!command" lines (readobj.c:58-61)

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
  void add(RObj&& o);
  void parse_stream(const std::string& text, const std::string& where, int depth);
};

}  // namespace rg
