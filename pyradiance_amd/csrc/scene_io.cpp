// Scene input: Radiance .rad text (common/readobj.c:28-190) and octree files
// (common/readoct.c:36-141, common/sceneio.c:21-109, common/portio.c:112-151).
//
// Only the object list is kept; the octree itself is skipped because the GPU
// path builds its own BVH from the primitives.
#include "scene_io.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
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
struct Lexer {
  const std::string& s;
  size_t p = 0;
  explicit Lexer(const std::string& t) : s(t) {}
  void skip_ws() { while (p < s.size() && isspace((unsigned char)s[p])) ++p; }
  bool eof() { skip_ws(); return p >= s.size(); }
  // fgetword(): whitespace-delimited word, quotes group words (common/fgetword.c)
  bool word(std::string& w) {
    skip_ws();
    if (p >= s.size()) return false;
    w.clear();
    char q = 0;
    while (p < s.size()) {
      char c = s[p];
      if (q) {
        if (c == q) { q = 0; ++p; continue; }
        w.push_back(c); ++p; continue;
      }
      if (isspace((unsigned char)c)) break;
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

bool parse_int(const std::string& w, long& v) {
  char* end = nullptr;
  v = strtol(w.c_str(), &end, 10);
  return end && *end == '\0' && !w.empty();
}
bool parse_flt(const std::string& w, double& v) {
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
      for (long i = 0; i < n; ++i) {
        double v;
        if (!lx.word(w) || !parse_flt(w, v)) throw RadError("(" + where + "): bad arguments for " + o.name);
        o.fargs.push_back(v);
      }
    }
    if (o.omod == -2) throw RadError("(" + where + "): inappropriate use of 'inherit' modifier");
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

void SceneDB::load_rad_file(const std::string& path) { parse_stream(slurp(path), path, 0); }

// ---- octree ---------------------------------------------------------------

namespace {
struct OctReader {
  const std::string& b;
  size_t p = 0;
  const std::string& fn;
  OctReader(const std::string& buf, const std::string& name) : b(buf), fn(name) {}
  [[noreturn]] void fail(const char* m) { throw RadError("octree \"" + fn + "\": " + m); }
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
}  // namespace

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
  // readscene(): type map, then objects until a -1 type byte (sceneio.c:90-109)
  std::vector<std::string> tmap;
  for (std::string s = rd.getstr(); !s.empty(); s = rd.getstr()) tmap.push_back(s);
  size_t base = objs.size();
  for (;;) {
    long ti = rd.getint(1);
    if (ti == -1) break;
    if (ti < 0 || ti >= (long)tmap.size()) throw RadError("octree \"" + path + "\": reference to unknown type");
    RObj o;
    o.type = tmap[ti];
    long m = rd.getint((int)objsize);
    o.omod = m == -1 ? -1 : (int)(m + (long)base);
    o.name = rd.getstr();
    long ns = rd.getint(2);
    for (long i = 0; i < ns; ++i) o.sargs.push_back(rd.getstr());
    long nf = rd.getint(2);
    for (long i = 0; i < nf; ++i) o.fargs.push_back(rd.getflt());
    if (!known_types().count(o.type)) throw RadError("octree \"" + path + "\": unknown object type \"" + o.type + "\"");
    add(std::move(o));
  }
  if ((long)(objs.size() - base) != nobj)
    throw RadError("octree \"" + path + "\": bad object count");
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
