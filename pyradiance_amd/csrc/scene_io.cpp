// Scene input: Radiance .rad text (common/readobj.c:28-190) and octree files
// (common/readoct.c:36-141, common/sceneio.c:21-109, common/portio.c:112-151).
//
// Only the object list is kept; the octree itself is skipped because the GPU
// path builds its own BVH from the primitives.
#include "scene_io.hpp"

#include <charconv>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <array>
#include <unordered_set>

namespace rg {

static const std::unordered_set<std::string>& surface_types() {
  static const std::unordered_set<std::string> s = {
      "polygon", "cone", "sphere", "ring", "cylinder", "cup", "bubble", "tube", "source"};
  return s;
}
static const std::unordered_set<std::string>& material_types() {
  // T_M entries of common/otypes.h:127-185
  static const std::unordered_set<std::string> s = {
      "plastic", "metal", "glass", "trans", "dielectric", "plastic2", "metal2", "trans2",
      "interface", "plasfunc", "metfunc", "glow", "light", "illum", "spotlight", "mist",
      "mirror", "transfunc", "BRTDfunc", "BSDF", "aBSDF", "WGMDfunc", "plasdata", "metdata",
      "transdata", "antimatter", "prism1", "prism2", "ashik2"};
  return s;
}
static const std::unordered_set<std::string>& known_types() {
  static const std::unordered_set<std::string> s = {
      "polygon", "cone", "sphere", "texfunc", "ring", "cylinder", "instance", "cup", "bubble",
      "tube", "mesh", "plastic", "metal", "glass", "trans", "dielectric", "plastic2", "metal2",
      "trans2", "interface", "plasfunc", "metfunc", "brightfunc", "brightdata", "brighttext",
      "colorpict", "glow", "source", "light", "illum", "spotlight", "mist", "mirror",
      "transfunc", "BRTDfunc", "BSDF", "aBSDF", "WGMDfunc", "plasdata", "metdata", "transdata",
      "colorfunc", "antimatter", "colordata", "colortext", "texdata", "mixfunc", "mixdata",
      "mixtext", "mixpict", "prism1", "prism2", "ashik2", "spectrum", "specfile", "specfunc",
      "specdata", "specpict", "alias"};
  return s;
}

bool SceneDB::is_surface(const std::string& t) { return surface_types().count(t) != 0; }
bool SceneDB::is_material(const std::string& t) { return material_types().count(t) != 0; }

int SceneDB::modifier(const std::string& name) const {
  auto it = modtab_.find(name);
  return it == modtab_.end() ? -1 : it->second;
}

void SceneDB::add(RObj&& o) {
  int idx = (int)objs.size();
  bool ismod = is_modifier(o.type);
  std::string nm = o.name;
  objs.push_back(std::move(o));
  if (ismod) modtab_[nm] = idx;
}

// ---- .rad text ------------------------------------------------------------

namespace {
// isspace() of the "C" locale
inline bool is_ws(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

struct Lexer {
  const std::string& s;
  size_t p = 0;
  explicit Lexer(const std::string& t) : s(t) {}
  void skip_ws() { while (p < s.size() && is_ws(s[p])) ++p; }
  bool eof() { skip_ws(); return p >= s.size(); }
  // fgetword(): whitespace-delimited word, quotes group words (common/fgetword.c)
  bool word(std::string& w) {
    skip_ws();
    if (p >= s.size()) return false;
    // common case: an unquoted word, copied in one piece
    size_t e = p;
    while (e < s.size() && !is_ws(s[e]) && s[e] != '"' && s[e] != '\'') ++e;
    if (e == s.size() || is_ws(s[e])) {
      w.assign(s, p, e - p);
      p = e;
      return true;
    }
    w.clear();
    char q = 0;
    while (p < s.size()) {
      char c = s[p];
      if (q) {
        if (c == q) { q = 0; ++p; continue; }
        w.push_back(c); ++p; continue;
      }
      if (is_ws(c)) break;
      if (c == '"' || c == '\'') { q = c; ++p; continue; }
      w.push_back(c); ++p;
    }
    return true;
  }
  std::string line() {
    size_t e = s.find('\n', p);
    if (e == std::string::npos) e = s.size();
    std::string l = s.substr(p, e - p);
    p = e;
    // continued lines ending in backslash (common/fgetline.c)
    while (!l.empty() && l.back() == '\\' && p < s.size()) {
      l.pop_back();
      ++p;
      size_t e2 = s.find('\n', p);
      if (e2 == std::string::npos) e2 = s.size();
      l += s.substr(p, e2 - p);
      p = e2;
    }
    return l;
  }
};

// Plain decimal words go through std::from_chars (correctly rounded, as
// glibc's strtod is, so the values are the same); anything else (a sign
// '+', hexadecimal, inf / nan) through the C library as before.
bool parse_int(const std::string& w, long& v) {
  const char* b = w.data();
  const char* e = b + w.size();
  auto r = std::from_chars(b, e, v, 10);
  if (r.ec == std::errc() && r.ptr == e && !w.empty()) return true;
  char* end = nullptr;
  v = strtol(w.c_str(), &end, 10);
  return end && *end == '\0' && !w.empty();
}
bool parse_flt(const std::string& w, double& v) {
  const char* b = w.data();
  const char* e = b + w.size();
  auto r = std::from_chars(b, e, v);
  if (r.ec == std::errc() && r.ptr == e && !w.empty()) return true;
  char* end = nullptr;
  v = strtod(w.c_str(), &end);
  return end && *end == '\0' && !w.empty();
}

std::string run_command(const std::string& cmd) {
  FILE* fp = popen(cmd.c_str(), "r");
  if (!fp) throw RadError("cannot execute \"" + cmd + "\"");
  std::string out;
  char buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, fp)) > 0) out.append(buf, n);
  int st = pclose(fp);
  if (st != 0) throw RadError("bad status from \"!" + cmd + "\"");
  return out;
}
}  // namespace

void SceneDB::parse_stream(const std::string& text, const std::string& where, int depth) {
  if (depth > 32) throw RadError("(" + where + "): command nesting too deep");
  Lexer lx(text);
  std::string w;
  while (!lx.eof()) {
    char c = lx.s[lx.p];
    if (c == '#') { lx.line(); continue; }
    if (c == '!') {
      ++lx.p;
      std::string cmd = lx.line();
      if (!allow_exec)
        throw RadError("(" + where + "): inline command \"!" + cmd +
                       "\" not executed (command execution disabled)");
      parse_stream(run_command(cmd), "!" + cmd, depth + 1);
      continue;
    }
    RObj o;
    std::string mod;
    if (!lx.word(mod)) break;
    if (!lx.word(o.type)) throw RadError("(" + where + "): unexpected EOF");
    if (!known_types().count(o.type))
      throw RadError("(" + where + "): unknown type \"" + o.type + "\"");
    if (!lx.word(o.name)) throw RadError("(" + where + "): unexpected EOF");
    if (mod == "void") o.omod = -1;
    else if (mod == "inherit") o.omod = -2;
    else {
      o.omod = modifier(mod);
      if (o.omod < 0) throw RadError("(" + where + "): undefined modifier \"" + mod + "\"");
    }
    if (o.type == "alias") {
      std::string ref;
      if (!lx.word(ref)) throw RadError("(" + where + "): unexpected EOF");
      int r = modifier(ref);
      if (r < 0) throw RadError("(" + where + "): bad reference \"" + ref + "\"");
      while (objs[r].type == "alias" && objs[r].sargs.empty() && objs[r].omod >= 0) r = objs[r].omod;
      if (o.omod == -2 || o.omod == objs[r].omod) o.omod = r;
      else o.sargs.push_back(ref);
    } else {
      long n;
      if (!lx.word(w) || !parse_int(w, n) || n < 0) throw RadError("(" + where + "): bad arguments for " + o.name);
      for (long i = 0; i < n; ++i) {
        if (!lx.word(w)) throw RadError("(" + where + "): bad arguments for " + o.name);
        o.sargs.push_back(w);
      }
      if (!lx.word(w) || !parse_int(w, n) || n != 0) throw RadError("(" + where + "): bad arguments for " + o.name);
      if (!lx.word(w) || !parse_int(w, n) || n < 0) throw RadError("(" + where + "): bad arguments for " + o.name);
      o.fargs.reserve((size_t)std::min<long>(n, 1L << 20));
      for (long i = 0; i < n; ++i) {
        double v;
        if (!lx.word(w) || !parse_flt(w, v)) throw RadError("(" + where + "): bad arguments for " + o.name);
        o.fargs.push_back(v);
      }
    }
    if (o.omod == -2) throw RadError("(" + where + "): inappropriate use of 'inherit' modifier");
    if (o.type == "mesh" || o.type == "instance") o.dir = cur_dir_;
    add(std::move(o));
  }
}

void SceneDB::load_rad_text(const std::string& text, const std::string& where) {
  parse_stream(text, where, 0);
}

static std::string slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw RadError("cannot open scene file \"" + path + "\"");
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

static std::string dir_of(const std::string& path) {
  size_t sl = path.rfind('/');
  return sl == std::string::npos ? std::string() : path.substr(0, sl + 1);
}

void SceneDB::load_rad_file(const std::string& path) {
  std::string keep = cur_dir_;
  cur_dir_ = dir_of(path);
  try {
    parse_stream(slurp(path), path, 0);
  } catch (...) {
    cur_dir_ = keep;
    throw;
  }
  cur_dir_ = keep;
}

// ---- octree ---------------------------------------------------------------

struct OctReader {
  const std::string& b;
  size_t p = 0;
  const std::string& fn;
  const char* what;
  OctReader(const std::string& buf, const std::string& name, const char* kind = "octree")
      : b(buf), fn(name), what(kind) {}
  [[noreturn]] void fail(const char* m) { throw RadError(std::string(what) + " \"" + fn + "\": " + m); }
  int getc() { return p < b.size() ? (unsigned char)b[p++] : -1; }
  // getint(): big-endian, sign-extended from the first byte (portio.c:112-133)
  long getint(int siz) {
    int c = getc();
    if (c < 0) fail("truncated octree");
    long r = c;
    if (c & 0x80) r |= -256L;
    while (--siz > 0) {
      if ((c = getc()) < 0) fail("truncated octree");
      r = (long)((unsigned long)r << 8) | c;
    }
    return r;
  }
  // getflt(): 31-bit mantissa and 8-bit exponent (portio.c:137-151)
  double getflt() {
    long l = getint(4);
    if (l == 0) { getc(); return 0.0; }
    double d = (l + .5 - (l < 0)) * (1. / 0x7fffffff);
    return ldexp(d, (int)getint(1));
  }
  std::string getstr() {
    size_t e = b.find('\0', p);
    if (e == std::string::npos) fail("truncated octree");
    std::string s = b.substr(p, e - p);
    p = e + 1;
    return s;
  }
  void skiptree(int objsize, int depth) {
    if (depth > 64) fail("damaged octree");
    int c = getc();
    switch (c) {
      case 0: return;                         // OT_EMPTY
      case 1: {                               // OT_FULL
        long n = getint(objsize);
        if (n < 0) fail("damaged octree");
        p += (size_t)n * objsize;
        if (p > b.size()) fail("truncated octree");
        return;
      }
      case 2:                                 // OT_TREE
        for (int i = 0; i < 8; ++i) skiptree(objsize, depth + 1);
        return;
      case -1: fail("truncated octree");
      default: fail("damaged octree");
    }
  }
};

// readscene(): type map, then objects until a -1 type byte (sceneio.c:90-109);
// modifier numbers are relative to the first object of the block.
void SceneDB::read_objects(OctReader& rd, int objsize, const std::string& dir) {
  std::vector<std::string> tmap;
  for (std::string s = rd.getstr(); !s.empty(); s = rd.getstr()) tmap.push_back(s);
  size_t base = objs.size();
  for (;;) {
    long ti = rd.getint(1);
    if (ti == -1) break;
    if (ti < 0 || ti >= (long)tmap.size()) rd.fail("reference to unknown type");
    RObj o;
    o.type = tmap[ti];
    long m = rd.getint(objsize);
    o.omod = m == -1 ? -1 : (int)(m + (long)base);
    o.name = rd.getstr();
    long ns = rd.getint(2);
    for (long i = 0; i < ns; ++i) o.sargs.push_back(rd.getstr());
    long nf = rd.getint(2);
    for (long i = 0; i < nf; ++i) o.fargs.push_back(rd.getflt());
    if (!known_types().count(o.type)) rd.fail(("unknown object type \"" + o.type + "\"").c_str());
    if (o.type == "mesh" || o.type == "instance") o.dir = dir;
    add(std::move(o));
  }
}

void SceneDB::load_oct(const std::string& path) {
  std::string buf = slurp(path);
  OctReader rd(buf, path);
  // header: text lines up to a blank line, must carry FORMAT=Radiance_octree
  size_t hend = buf.find("\n\n");
  if (buf.compare(0, 10, "#?RADIANCE") != 0 || hend == std::string::npos)
    throw RadError("\"" + path + "\": not an octree");
  std::string header = buf.substr(0, hend);
  if (header.find("FORMAT=Radiance_octree") == std::string::npos)
    throw RadError("\"" + path + "\": not an octree");
  rd.p = hend + 2;
  const long OCTMAGIC = 4 * 8 + 251;          // common/octree.h:74-75
  long objsize = rd.getint(2) - OCTMAGIC;
  if (objsize <= 0 || objsize > 8) throw RadError("\"" + path + "\": incompatible octree format");
  for (int i = 0; i < 3; ++i) cuorg[i] = atof(rd.getstr().c_str());
  cusize = atof(rd.getstr().c_str());
  std::vector<std::string> files;
  for (std::string s = rd.getstr(); !s.empty(); s = rd.getstr()) files.push_back(s);
  long nobj = rd.getint((int)objsize);
  if (!files.empty()) {
    std::string dir;
    size_t sl = path.rfind('/');
    if (sl != std::string::npos) dir = path.substr(0, sl + 1);
    size_t before = objs.size();
    for (auto& f : files) {
      if (f[0] == '!') {
        if (!allow_exec) throw RadError("octree \"" + path + "\" names command \"" + f + "\" (command execution disabled)");
        parse_stream(run_command(f.substr(1)), f, 1);
        continue;
      }
      // relative names are relative to where oconv ran; try the current
      // directory, then the octree's directory, then the bare file name there
      std::vector<std::string> cands = {f};
      if (f[0] != '/' && !dir.empty()) {
        cands.push_back(dir + f);
        size_t fs = f.rfind('/');
        if (fs != std::string::npos) cands.push_back(dir + f.substr(fs + 1));
      }
      std::string use;
      for (auto& cnd : cands) {
        std::ifstream probe(cnd);
        if (probe.good()) { use = cnd; break; }
      }
      if (use.empty()) throw RadError("octree \"" + path + "\": cannot open scene file \"" + f + "\"");
      load_rad_file(use);
    }
    if ((long)(objs.size() - before) != nobj)
      throw RadError("octree \"" + path + "\": bad object count; octree stale?");
    return;
  }
  rd.skiptree((int)objsize, 0);
  size_t base = objs.size();
  read_objects(rd, (int)objsize, dir_of(path));
  if ((long)(objs.size() - base) != nobj)
    throw RadError("octree \"" + path + "\": bad object count");
}

// ---- compiled meshes (common/readmesh.c:190-300) ------------------------------

// decodedir() of common/dircode.c:83-106: two 13-bit components, the axis
// they belong to and three sign bits; 0 is "no normal".
static void decode_dir(double dv[3], int32_t dc) {
  static const int itab[4][3] = {{1, 0, 2}, {0, 1, 2}, {1, 2, 0}, {0, 2, 1}};
  const double DCSCALE = 11584.5;
  const int ndx = ((dc & 020) != 0) << 1 | ((dc & 010) != 0);
  double d1 = (dc >> 5 & 0x1fff) * (1. / DCSCALE);
  double d2 = (dc >> 18 & 0x1fff) * (1. / DCSCALE);
  double der = sqrt(1. - d1 * d1 - d2 * d2);
  dv[itab[ndx][0]] = d1;
  dv[itab[ndx][1]] = d2;
  dv[itab[ndx][2]] = der;
  if (dc & 01) dv[0] = -dv[0];
  if (dc & 02) dv[1] = -dv[1];
  if (dc & 04) dv[2] = -dv[2];
}

int SceneDB::load_mesh(const std::string& path) {
  auto it = mesh_files_.find(path);
  if (it != mesh_files_.end()) return it->second;
  std::string buf = slurp(path);
  OctReader rd(buf, path, "mesh");
  size_t hend = buf.find("\n\n");
  if (buf.compare(0, 10, "#?RADIANCE") != 0 || hend == std::string::npos ||
      buf.substr(0, hend).find("FORMAT=Radiance_tmesh") == std::string::npos)
    throw RadError("\"" + path + "\": not a compiled mesh");
  rd.p = hend + 2;
  const long MESHMAGIC = 1 * 8 + 311;         // common/mesh.h:128-129
  long objsize = rd.getint(2) - MESHMAGIC;
  if (objsize <= 0 || objsize > 8) throw RadError("\"" + path + "\": incompatible mesh format");
  double org[3], size;
  for (int i = 0; i < 3; ++i) org[i] = atof(rd.getstr().c_str());
  size = atof(rd.getstr().c_str());
  for (int i = 0; i < 4; ++i) rd.getflt();    // (u,v) limits: no patterns here
  rd.skiptree((int)objsize, 0);
  MeshData M;
  M.path = path;
  M.mat0 = (int)objs.size();
  read_objects(rd, (int)objsize, dir_of(path));
  M.nmats = (int)objs.size() - M.mat0;
  long npatch = rd.getint(4);
  if (npatch < 0 || npatch > (1L << 22)) rd.fail("bad patch count");
  struct J1 { int32_t v1j; int v2, v3, mat; };
  struct J2 { int32_t v1j, v2j; int v3, mat; };
  std::vector<int64_t> base(npatch + 1, 0);
  std::vector<std::vector<std::array<int, 4>>> local(npatch);
  std::vector<std::vector<J1>> j1(npatch);
  std::vector<std::vector<J2>> j2(npatch);
  const double vres = (1. / 4294967296.) * size;
  for (long pn = 0; pn < npatch; ++pn) {
    int flags = (int)rd.getint(1);
    if (!(flags & 1) || (flags & ~7)) rd.fail("bad patch flags");
    int nv = (int)rd.getint(2);
    if (nv <= 0 || nv > 256) rd.fail("bad number of patch vertices");
    base[pn + 1] = base[pn] + nv;
    size_t v0 = M.has_n.size();
    M.xyz.resize(3 * (v0 + nv));
    M.nrm.resize(3 * (v0 + nv), 0.0);
    M.has_n.resize(v0 + nv, 0);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < 3; ++j)
        M.xyz[3 * (v0 + i) + j] = org[j] + ((uint32_t)rd.getint(4) + .5) * vres;
    if (flags & 2)
      for (int i = 0; i < nv; ++i) {
        int32_t code = (int32_t)rd.getint(4);
        if (code) {
          decode_dir(&M.nrm[3 * (v0 + i)], code);
          M.has_n[v0 + i] = 1;
        }
      }
    if (flags & 4)
      for (int i = 0; i < 2 * nv; ++i) rd.getint(4);
    int ntris = (int)rd.getint(2);
    if (ntris < 0 || ntris > 512) rd.fail("bad number of local triangles");
    auto& lt = local[pn];
    lt.resize(ntris);
    for (int i = 0; i < ntris; ++i)
      for (int k = 0; k < 3; ++k) lt[i][k] = (int)(rd.getint(1) & 0xff);
    if (rd.getint(2) > 1) {
      for (int i = 0; i < ntris; ++i) lt[i][3] = (int)rd.getint(2);
    } else {
      int sole = (int)rd.getint(2);
      for (int i = 0; i < ntris; ++i) lt[i][3] = sole;
    }
    int n1 = (int)rd.getint(2);
    if (n1 < 0 || n1 > 256) rd.fail("bad number of joiner triangles");
    for (int i = 0; i < n1; ++i) {
      J1 t;
      t.v1j = (int32_t)rd.getint(4);
      t.v2 = (int)(rd.getint(1) & 0xff);
      t.v3 = (int)(rd.getint(1) & 0xff);
      t.mat = (int)rd.getint(2);
      j1[pn].push_back(t);
    }
    int n2 = (int)rd.getint(2);
    if (n2 < 0 || n2 > 256) rd.fail("bad number of double joiner triangles");
    for (int i = 0; i < n2; ++i) {
      J2 t;
      t.v1j = (int32_t)rd.getint(4);
      t.v2j = (int32_t)rd.getint(4);
      t.v3 = (int)(rd.getint(1) & 0xff);
      t.mat = (int)rd.getint(2);
      j2[pn].push_back(t);
    }
  }
  // vertex id (patch << 8 | local) -> flat index; flat order follows the ids
  auto flat = [&](int64_t id) -> int32_t {
    int64_t pn = id >> 8, k = id & 0xff;
    if (pn < 0 || pn >= npatch || base[pn] + k >= base[pn + 1]) rd.fail("bad vertex reference");
    return (int32_t)(base[pn] + k);
  };
  auto matof = [&](int m) { return m == -1 ? -1 : M.mat0 + m; };
  for (long pn = 0; pn < npatch; ++pn) {
    for (auto& t : local[pn])
      M.tris.push_back({{flat(pn << 8 | t[0]), flat(pn << 8 | t[1]), flat(pn << 8 | t[2])}, matof(t[3])});
    for (auto& t : j1[pn])
      M.tris.push_back({{flat(t.v1j), flat(pn << 8 | t.v2), flat(pn << 8 | t.v3)}, matof(t.mat)});
    for (auto& t : j2[pn])
      M.tris.push_back({{flat(t.v1j), flat(t.v2j), flat(pn << 8 | t.v3)}, matof(t.mat)});
  }
  for (auto& t : M.tris)
    if (t.mat >= 0 && t.mat >= M.mat0 + M.nmats) rd.fail("material index out of range");
  int idx = (int)meshes.size();
  meshes.push_back(std::move(M));
  mesh_files_[path] = idx;
  return idx;
}

// The file a mesh or instance object names: as given (relative to the
// current directory, as the reference's library path starts with "."), then
// next to the file that defined the object.
std::string SceneDB::find_file(const RObj& o, const char* what) const {
  if (o.sargs.empty()) throw RadError(o.type + " \"" + o.name + "\": bad # of arguments");
  const std::string& f = o.sargs[0];
  std::vector<std::string> cands = {f};
  if (f[0] != '/' && !o.dir.empty()) cands.push_back(o.dir + f);
  for (auto& c : cands) {
    std::ifstream probe(c);
    if (probe.good()) return c;
  }
  throw RadError(std::string("cannot find ") + what + " file \"" + f + "\"");
}

void SceneDB::resolve_meshes() {
  for (int i = 0; i < (int)objs.size(); ++i) {
    if (objs[i].type != "mesh" || obj_mesh.count(i)) continue;
    const RObj o = objs[i];
    obj_mesh[i] = load_mesh(find_file(o, "mesh"));
  }
}

void SceneDB::resolve_instances() {
  // objs grows while we walk it: nested instances are reached too
  for (int i = 0; i < (int)objs.size(); ++i) {
    if (objs[i].type != "instance" || obj_inst.count(i)) continue;
    const RObj o = objs[i];
    std::string use = find_file(o, "octree");
    auto it = inst_files_.find(use);
    if (it != inst_files_.end()) {
      obj_inst[i] = it->second;
      continue;
    }
    if (inst_files_.size() > 4096) throw RadError("too many instanced octrees");
    double keep_org[3] = {cuorg[0], cuorg[1], cuorg[2]}, keep_size = cusize;
    int first = (int)objs.size();
    load_oct(use);
    for (int k = 0; k < 3; ++k) cuorg[k] = keep_org[k];
    cusize = keep_size;
    int b = (int)inst_blocks.size();
    inst_blocks.push_back({first, (int)objs.size() - first});
    inst_files_[use] = b;
    obj_inst[i] = b;
  }
  hidden.assign(objs.size(), 0);
  for (const auto& b : inst_blocks)
    for (int k = b.first; k < b.first + b.count; ++k) hidden[k] = 1;
}

void SceneDB::load_any(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw RadError("cannot open \"" + path + "\"");
  char head[16] = {0};
  f.read(head, 10);
  std::string h(head, (size_t)f.gcount());
  if (h == "#?RADIANCE") {
    std::string rest((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t e = rest.find("\n\n");
    if (e != std::string::npos && rest.substr(0, e).find("FORMAT=Radiance_octree") != std::string::npos) {
      load_oct(path);
      return;
    }
  }
  load_rad_file(path);
}

}  // namespace rg
