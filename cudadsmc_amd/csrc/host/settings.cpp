// Settings: defaults, config.xml / gases.xml reader and command line.
//
// Mirrors settings_t (include/sim_settings.h:32-69, src/sim_settings.cpp:
// 34-186) and main's particle count (kernels/dsmc.cu:324-330). The XML
// reader is a small recursive-descent parser for the element/attribute/text
// subset these files use (no DTDs or entities beyond the five predefined).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "dsmc/dsmc_c.h"
#include "host_internal.h"

namespace dsmc_host {

// ---------------------------------------------------------------- XML DOM
struct XmlNode {
  std::string name;
  std::string text;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<XmlNode>> children;
  const XmlNode* child(const char* n) const {
    for (auto& c : children)
      if (c->name == n) return c.get();
    return nullptr;
  }
  const char* attr(const char* n) const {
    for (auto& a : attrs)
      if (a.first == n) return a.second.c_str();
    return nullptr;
  }
};

class XmlParser {
 public:
  explicit XmlParser(const std::string& s) : s_(s), i_(0) {}
  std::unique_ptr<XmlNode> parse_document(std::string& err) {
    skip_misc();
    if (i_ >= s_.size() || s_[i_] != '<') {
      err = "no root element";
      return nullptr;
    }
    auto n = parse_element(err);
    return n;
  }

 private:
  const std::string& s_;
  size_t i_;

  void skip_ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
  }
  bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
  void skip_until(const char* p) {
    size_t k = s_.find(p, i_);
    i_ = (k == std::string::npos) ? s_.size() : k + std::strlen(p);
  }
  // comments, processing instructions, doctype
  void skip_misc() {
    for (;;) {
      skip_ws();
      if (starts("<?")) skip_until("?>");
      else if (starts("<!--")) skip_until("-->");
      else if (starts("<!")) skip_until(">");
      else return;
    }
  }
  static std::string decode(const std::string& t) {
    std::string o;
    for (size_t k = 0; k < t.size(); ++k) {
      if (t[k] == '&') {
        size_t e = t.find(';', k);
        if (e != std::string::npos) {
          std::string ent = t.substr(k + 1, e - k - 1);
          const char* rep = ent == "lt" ? "<" : ent == "gt" ? ">" : ent == "amp" ? "&" : ent == "quot" ? "\"" :
                            ent == "apos" ? "'" : nullptr;
          if (rep) {
            o += rep;
            k = e;
            continue;
          }
        }
      }
      o += t[k];
    }
    return o;
  }
  std::string parse_name() {
    size_t b = i_;
    while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == '-' || s_[i_] == ':' ||
                              s_[i_] == '.'))
      ++i_;
    return s_.substr(b, i_ - b);
  }
  std::unique_ptr<XmlNode> parse_element(std::string& err) {
    auto n = std::make_unique<XmlNode>();
    ++i_;  // '<'
    n->name = parse_name();
    if (n->name.empty()) {
      err = "bad element name";
      return nullptr;
    }
    for (;;) {
      skip_ws();
      if (i_ >= s_.size()) {
        err = "unterminated element <" + n->name + ">";
        return nullptr;
      }
      if (starts("/>")) {
        i_ += 2;
        return n;
      }
      if (s_[i_] == '>') {
        ++i_;
        break;
      }
      std::string an = parse_name();
      skip_ws();
      if (an.empty() || i_ >= s_.size() || s_[i_] != '=') {
        err = "bad attribute in <" + n->name + ">";
        return nullptr;
      }
      ++i_;
      skip_ws();
      if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) {
        err = "unquoted attribute in <" + n->name + ">";
        return nullptr;
      }
      char q = s_[i_++];
      size_t e = s_.find(q, i_);
      if (e == std::string::npos) {
        err = "unterminated attribute";
        return nullptr;
      }
      n->attrs.push_back({an, decode(s_.substr(i_, e - i_))});
      i_ = e + 1;
    }
    // content
    for (;;) {
      if (i_ >= s_.size()) {
        err = "missing </" + n->name + ">";
        return nullptr;
      }
      if (starts("<!--")) {
        skip_until("-->");
        continue;
      }
      if (starts("</")) {
        i_ += 2;
        std::string cn = parse_name();
        skip_ws();
        if (cn != n->name || i_ >= s_.size() || s_[i_] != '>') {
          err = "mismatched </" + cn + "> for <" + n->name + ">";
          return nullptr;
        }
        ++i_;
        return n;
      }
      if (s_[i_] == '<') {
        auto c = parse_element(err);
        if (!c) return nullptr;
        n->children.push_back(std::move(c));
        continue;
      }
      size_t e = s_.find('<', i_);
      if (e == std::string::npos) e = s_.size();
      n->text += decode(s_.substr(i_, e - i_));
      i_ = e;
    }
  }
};

static bool read_text(const std::string& path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

static std::string dir_of(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

static void trim(std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
  s = (b == std::string::npos) ? std::string() : s.substr(b, e - b + 1);
}

static void cell_size_from_grid(dsmc_settings* s) {
  for (int a = 0; a < 3; ++a) s->cell_size[a] = (s->bmax[a] - s->bmin[a]) / (float)s->grid_dim[a];
}

}  // namespace dsmc_host

using namespace dsmc_host;

extern "C" {

void dsmc_settings_default(dsmc_settings* s) {
  std::memset(s, 0, sizeof(*s));
  // settings_t::settings_t (src/sim_settings.cpp:34-44) with include/dsmc.h defaults
  s->benchmark = 0;
  s->cpu = 0;
  s->geometry = 1;
  s->mt = 0;
  s->num_particles = 4096u * 256u;  // SAMPLES_COUNT
  s->max_run_time = -1.0f;
  s->density = 1.0e20f;   // DSMC_DENSITY
  s->fnum = 0.2e14f;      // DSMC_FNUM
  s->temp = 300.0f;       // DSMC_T
  s->grid_dim[0] = 30;    // GRID_DIM_X
  s->grid_dim[1] = 18;
  s->grid_dim[2] = 18;
  s->dt = 2e-6f;          // DSMC_DT
  s->periodic_condition = 'n';
  s->bmin[0] = 0.0f; s->bmin[1] = 0.0f; s->bmin[2] = 0.0f;
  s->bmax[0] = 0.3f; s->bmax[1] = 0.18f; s->bmax[2] = 0.18f;
  cell_size_from_grid(s);
  s->molmass = 5.0e-26f;  // DSMC_MOL_MASS
  s->diameter = 3.5e-10f; // DSMC_DIAM
  s->Tref = 273.0f;       // DSMC_T0
  std::snprintf(s->gas_name, sizeof(s->gas_name), "%s", "default");
  // extensions
  s->mode = DSMC_MODE_REFERENCE;
  s->omega = 0.5f;
  s->wall_temp = 0.0f;
  for (int f = 0; f < 6; ++f) s->face_bc[f] = DSMC_FACE_SPECULAR;
  s->seed = 0;
  s->sample_every = 0;
}

int dsmc_settings_read_file(dsmc_settings* s, const char* path) {
  if (!s || !path) return set_global_error(DSMC_ERR_INVALID, "null argument");
  std::string text;
  if (!read_text(path, text)) return set_global_error(DSMC_ERR_IO, std::string("cannot open config ") + path);
  std::string err;
  XmlParser px(text);
  auto root = px.parse_document(err);
  if (!root) return set_global_error(DSMC_ERR_IO, std::string(path) + ": " + err);
  auto need = [&](const XmlNode* n, const char* what) -> bool {
    if (!n) {
      set_global_error(DSMC_ERR_IO, std::string(path) + ": missing <" + what + ">");
      return false;
    }
    return true;
  };
  const XmlNode* grid = root->child("grid");
  if (!need(grid, "grid")) return DSMC_ERR_IO;
  const XmlNode *gmin = grid->child("min"), *gmax = grid->child("max"), *gdim = grid->child("dim");
  if (!need(gmin, "min") || !need(gmax, "max") || !need(gdim, "dim")) return DSMC_ERR_IO;
  if (std::sscanf(gmin->text.c_str(), "%f %f %f", &s->bmin[0], &s->bmin[1], &s->bmin[2]) != 3 ||
      std::sscanf(gmax->text.c_str(), "%f %f %f", &s->bmax[0], &s->bmax[1], &s->bmax[2]) != 3 ||
      std::sscanf(gdim->text.c_str(), "%d %d %d", &s->grid_dim[0], &s->grid_dim[1], &s->grid_dim[2]) != 3)
    return set_global_error(DSMC_ERR_IO, std::string(path) + ": bad <grid>");
  for (int a = 0; a < 3; ++a)
    if (s->grid_dim[a] <= 0) return set_global_error(DSMC_ERR_INVALID, "grid dimensions must be positive");
  cell_size_from_grid(s);
  const XmlNode *dn = root->child("density"), *fn = root->child("fnum"), *tn = root->child("temp");
  if (!need(dn, "density") || !need(fn, "fnum") || !need(tn, "temp")) return DSMC_ERR_IO;
  std::sscanf(dn->text.c_str(), "%f", &s->density);
  std::sscanf(fn->text.c_str(), "%f", &s->fnum);
  std::sscanf(tn->text.c_str(), "%f", &s->temp);
  if (const XmlNode* sv = root->child("streamvel"))
    std::sscanf(sv->text.c_str(), "%f %f %f", &s->stream_velocity[0], &s->stream_velocity[1],
                &s->stream_velocity[2]);
  if (const XmlNode* pn = root->child("periodic")) {
    std::string v = pn->text;
    trim(v);
    s->periodic_condition = v.empty() ? '\0' : v[0];
  }
  // ---- extensions (absent from the reference's files)
  if (const XmlNode* m = root->child("mode")) {
    std::string v = m->text;
    trim(v);
    s->mode = (v == "physics") ? DSMC_MODE_PHYSICS : DSMC_MODE_REFERENCE;
  }
  if (const XmlNode* n = root->child("omega")) std::sscanf(n->text.c_str(), "%f", &s->omega);
  if (const XmlNode* n = root->child("wall_temp")) std::sscanf(n->text.c_str(), "%f", &s->wall_temp);
  if (const XmlNode* n = root->child("particles")) std::sscanf(n->text.c_str(), "%u", &s->target_particles);
  if (const XmlNode* n = root->child("seed")) std::sscanf(n->text.c_str(), "%u", &s->seed);
  if (const XmlNode* n = root->child("sample_every")) std::sscanf(n->text.c_str(), "%u", &s->sample_every);
  if (const XmlNode* n = root->child("dt")) std::sscanf(n->text.c_str(), "%f", &s->dt);
  if (const XmlNode* n = root->child("boundary")) {
    // <boundary>freestream|specular|periodic ...</boundary>: one word for all
    // faces or six words (-x +x -y +y -z +z)
    std::istringstream is(n->text);
    std::vector<std::string> w;
    std::string t;
    while (is >> t) w.push_back(t);
    auto kind = [](const std::string& k) {
      return k == "freestream" ? DSMC_FACE_FREESTREAM : k == "periodic" ? DSMC_FACE_PERIODIC : DSMC_FACE_SPECULAR;
    };
    for (int f = 0; f < 6; ++f)
      if (!w.empty()) s->face_bc[f] = kind(w.size() == 6 ? w[f] : w[0]);
  }
  // ---- gas
  const XmlNode* gas = root->child("gas");
  if (!need(gas, "gas")) return DSMC_ERR_IO;
  const char* gfile = gas->attr("file");
  if (!gfile) return set_global_error(DSMC_ERR_IO, std::string(path) + ": <gas> without file attribute");
  std::string gname = gas->text;
  trim(gname);
  std::snprintf(s->gas_name, sizeof(s->gas_name), "%s", gname.c_str());
  if (const XmlNode* o = gas->child("omega")) std::sscanf(o->text.c_str(), "%f", &s->omega);
  std::string gtext;
  std::string gpath = gfile;
  if (!read_text(gpath, gtext)) {  // reference: relative to the working directory
    gpath = dir_of(path) + "/" + gfile;
    if (!read_text(gpath, gtext)) return set_global_error(DSMC_ERR_IO, std::string("cannot open gases file ") + gfile);
  }
  XmlParser gp(gtext);
  auto groot = gp.parse_document(err);
  if (!groot) return set_global_error(DSMC_ERR_IO, gpath + ": " + err);
  const XmlNode* found = nullptr;
  for (auto& c : groot->children) {
    if (c->name != "gas") continue;
    const char* nm = c->attr("name");
    if (nm && gname == nm) {
      found = c.get();
      break;
    }
  }
  if (!found) return set_global_error(DSMC_ERR_IO, "gas '" + gname + "' not found in " + gpath);
  const XmlNode *ms = found->child("mass"), *dm = found->child("diam"), *tr = found->child("temp");
  if (ms) std::sscanf(ms->text.c_str(), "%f", &s->molmass);
  if (dm) std::sscanf(dm->text.c_str(), "%f", &s->diameter);
  if (tr) std::sscanf(tr->text.c_str(), "%f", &s->Tref);
  if (const XmlNode* o = found->child("omega")) std::sscanf(o->text.c_str(), "%f", &s->omega);
  return DSMC_OK;
}

int dsmc_settings_read_cmdline(dsmc_settings* s, int argc, const char* const* argv) {
  if (!s) return set_global_error(DSMC_ERR_INVALID, "null settings");
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    bool has_next = i + 1 < argc;
    if (!std::strcmp(a, "-f")) {
      if (!has_next) return set_global_error(DSMC_ERR_INVALID, "-f needs a file");
      int rc = dsmc_settings_read_file(s, argv[i + 1]);
      if (rc) return rc;
      continue;
    }
    if (!std::strcmp(a, "-benchmark")) { s->benchmark = 1; continue; }
    if (!std::strcmp(a, "-cpu")) { s->cpu = 1; continue; }
    if (!std::strcmp(a, "-no-geometry")) { s->geometry = 0; continue; }
    if (!std::strcmp(a, "-mt")) { s->mt = 1; continue; }
    if (!std::strcmp(a, "-density")) {
      if (!has_next) return set_global_error(DSMC_ERR_INVALID, "-density needs a value");
      s->density = (float)std::atof(argv[i + 1]);
      continue;
    }
    if (!std::strcmp(a, "-dt")) {
      if (!has_next) return set_global_error(DSMC_ERR_INVALID, "-dt needs a value");
      s->dt = (float)std::atof(argv[i + 1]);
      continue;
    }
    if (!std::strcmp(a, "-run-for")) {
      if (!has_next) return set_global_error(DSMC_ERR_INVALID, "-run-for needs a value");
      s->max_run_time = (float)std::atof(argv[i + 1]);
      continue;
    }
  }
  return DSMC_OK;
}

int dsmc_settings_derive_particles(dsmc_settings* s) {
  if (!s) return set_global_error(DSMC_ERR_INVALID, "null settings");
  // kernels/dsmc.cu:324-330, FP32 as in the reference
  float cell_volume = s->cell_size[0] * s->cell_size[1] * s->cell_size[2];
  float ppc_f = s->density * cell_volume / s->fnum;
  uint32_t ppc = (ppc_f > 0.0f && ppc_f < 4294967296.0f) ? (uint32_t)ppc_f : 0u;
  uint64_t cells = (uint64_t)s->grid_dim[0] * (uint64_t)s->grid_dim[1] * (uint64_t)s->grid_dim[2];
  if ((uint64_t)ppc * cells > 0xFFFFFFFFull) return set_global_error(DSMC_ERR_CAPACITY, "particle count overflows");
  s->part_per_cell = ppc;
  s->num_particles = (uint32_t)((uint64_t)ppc * cells);
  return DSMC_OK;
}

}  // extern "C"
