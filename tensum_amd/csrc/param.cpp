// param.cpp -- param.in reader (grammar of src/reader.h, sections and checks
// of src/parameter.cc).
#include "param.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace tsm {

namespace {

// Token reader with the reference's comment rules (src/reader.h:50-89).
class Reader {
 public:
  explicit Reader(std::string text) : s_(std::move(text)) {}
  void skip() {
    for (;;) {
      while (p_ < s_.size() && (std::isspace((unsigned char)s_[p_]) || s_[p_] == ',' || s_[p_] == ';')) ++p_;
      if (p_ >= s_.size()) return;
      char c = s_[p_];
      if (c == '#' || c == '%' || (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/')) {
        while (p_ < s_.size() && s_[p_] != '\n') ++p_;
        continue;
      }
      if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '*') {
        size_t e = s_.find("*/", p_ + 2);
        p_ = (e == std::string::npos) ? s_.size() : e + 2;
        continue;
      }
      return;
    }
  }
  bool word(std::string& w) {
    skip();
    size_t b = p_;
    while (p_ < s_.size() && !std::isspace((unsigned char)s_[p_])) ++p_;
    w = s_.substr(b, p_ - b);
    return !w.empty();
  }
  bool num(double& d) {
    std::string w;
    if (!word(w)) return false;
    char* e = nullptr;
    d = std::strtod(w.c_str(), &e);
    return e && *e == '\0';
  }
  bool integer(long& v) {
    std::string w;
    if (!word(w)) return false;
    char* e = nullptr;
    v = std::strtol(w.c_str(), &e, 10);
    return e && *e == '\0';
  }
  void line(std::string& l) {
    skip();
    size_t e = s_.find('\n', p_);
    if (e == std::string::npos) e = s_.size();
    l = s_.substr(p_, e - p_);
    p_ = e;
    while (!l.empty() && std::isspace((unsigned char)l.back())) l.pop_back();
  }
  bool peek_is(const char* tok) {
    size_t save = p_;
    std::string w;
    word(w);
    if (w == tok) return true;
    p_ = save;
    return false;
  }
  bool eos() { return peek_is("}"); }
  bool bos() { return peek_is("{"); }
  bool entry(const std::string& want, std::string& err) {
    std::string w;
    word(w);
    if (w != want) { err = "Expecting = " + want + ", but found = " + w; return false; }
    return true;
  }
  bool begin(const std::string& name, std::string& err) { return entry(name, err) && entry("{", err); }

 private:
  std::string s_;
  size_t p_ = 0;
};

#define NEED(cond, msg) do { if (!(cond)) { err = (msg); return false; } } while (0)
#define ENTRY(name) do { if (!R.entry(name, err)) return false; } while (0)
#define BEGIN(name) do { if (!R.begin(name, err)) return false; } while (0)

bool yesno(Reader& R, const char* what, bool& out, std::string& err) {
  std::string w;
  R.word(w);
  if (w == "yes") out = true;
  else if (w == "no") out = false;
  else { err = std::string("unknown option ") + w + " for " + what + ". Possible options: yes, no"; return false; }
  return true;
}

}  // namespace

bool read_settings(const std::string& file, Settings& S, std::string& err) {
  std::ifstream in(file);
  if (!in) { err = "cannot open " + file; return false; }
  std::stringstream ss;
  ss << in.rdbuf();
  Reader R(ss.str());
  tsm_params& p = S.p;
  p = tsm_params{};
  p.nd_active = p.visc_active = p.heat_active = 1;
  std::string w;

  // grid (src/parameter.cc:38-56)
  BEGIN("grid");
  ENTRY("cell");
  R.word(w);
  if (w == "median") p.cell_type = 0;
  else if (w == "voronoi") p.cell_type = 1;
  else { err = "read_grid: unknown cell type " + w; return false; }
  ENTRY("}");

  // numeric (src/parameter.cc:61-258)
  BEGIN("numeric");
  ENTRY("time_mode");
  R.word(S.time_mode);
  NEED(S.time_mode == "steady" || S.time_mode == "unsteady", "read_numeric: unknown time_mode " + S.time_mode);
  ENTRY("time_scheme");
  R.word(S.time_scheme);
  NEED(S.time_scheme == "heuns" || S.time_scheme == "ssprk3" || S.time_scheme == "rk4",
       "read_numeric: unknown time_scheme " + S.time_scheme);
  ENTRY("time_step");
  NEED(R.num(p.time_step) && p.time_step >= 0.0, "time_step must be >= 0");
  ENTRY("cfl");
  NEED(R.num(p.cfl) && p.cfl >= 0.0, "cfl must be >= 0");
  ENTRY("max_iter");
  long mi = 0;
  NEED(R.integer(mi) && mi > 0, "max_iter must be > 0");
  p.max_iter = mi;
  ENTRY("final_time");
  NEED(R.num(p.final_time) && p.final_time >= 0.0, "final_time must be >= 0");
  ENTRY("min_residue");
  NEED(R.num(p.min_residue) && p.min_residue >= 0.0, "min_residue must be >= 0");
  BEGIN("reconstruct");
  ENTRY("method");
  R.word(w);
  if (w == "first") p.reconstruct = TSM_REC_FIRST;
  else if (w == "second") p.reconstruct = TSM_REC_SECOND;
  else if (w == "minmod") p.reconstruct = TSM_REC_MINMOD;
  else if (w == "tvb_minmod") {
    p.reconstruct = TSM_REC_TVB_MINMOD;
    ENTRY("M_value");
    NEED(R.num(p.tvb_M) && p.tvb_M > 0.0, "M_value must be > 0");
  } else if (w == "eno2") p.reconstruct = TSM_REC_ENO2;
  else if (w == "van_albada") p.reconstruct = TSM_REC_VAN_ALBADA;
  else { err = "read_numeric: unknown reconstruction scheme " + w; return false; }
  ENTRY("}");
  ENTRY("bc_scheme");
  R.word(w);
  NEED(w == "weak" || w == "strong", "read_numeric: unknown bc scheme " + w);
  p.strong_bc = w == "strong";
  ENTRY("liou_fix");
  bool liou = false;
  if (!yesno(R, "liou_fix", liou, err)) return false;
  p.liou_fix = liou;
  BEGIN("sample_list");
  BEGIN("groups");
  while (!R.eos()) {
    long a, b;
    NEED(R.integer(a) && R.integer(b), "bad sample group");
    for (long i = a; i <= b; ++i) S.sample_list.push_back((int)i);
  }
  BEGIN("free_list");
  while (!R.eos()) {
    long a;
    NEED(R.integer(a), "bad free_list entry");
    S.sample_list.push_back((int)a);
  }
  ENTRY("}");
  ENTRY("rnd_file_loc");
  R.word(S.rnd_file_loc);
  ENTRY("rnd_per_sample");
  long rps = 0;
  NEED(R.integer(rps) && rps >= 0, "bad rnd_per_sample");
  S.rnd_per_sample = (unsigned)rps;
  ENTRY("}");
  if (S.sample_list.empty()) S.sample_list.push_back(0);
  std::sort(S.sample_list.begin(), S.sample_list.end());
  S.sample_list.erase(std::unique(S.sample_list.begin(), S.sample_list.end()), S.sample_list.end());
  NEED(!(p.time_step == 0.0 && p.cfl == 0.0), "Either time_step or cfl must be non-zero");
  NEED(!(p.time_step > 0.0 && p.cfl > 0.0), "You have specified both time_step and cfl. Specify only one of them, and set the other to zero");
  p.unsteady = S.time_mode == "unsteady";
  if (!p.unsteady) p.final_time = 1.0e20;
  else NEED(p.final_time > 0.0, "read_numeric: final_time must be positive for unsteady flow");
  p.time_scheme = S.time_scheme == "rk4" ? 1 : (S.time_scheme == "ssprk3" ? 0 : 2);

  // material (src/parameter.cc:263-461)
  BEGIN("material");
  ENTRY("gamma");
  NEED(R.num(p.gamma) && p.gamma > 1.0, "gamma must be > 1");
  ENTRY("gas_const");
  NEED(R.num(p.gas_const) && p.gas_const > 0.0, "gas_const must be > 0");
  ENTRY("model");
  R.word(w);
  if (w == "euler") p.model = 0;
  else if (w == "ns") p.model = 1;
  else { err = "read_material: unknown flow model " + w; return false; }
  BEGIN("viscosity");
  ENTRY("model");
  R.word(w);
  if (w == "constant") {
    p.mu_model = 0;
    ENTRY("mu_ref");
    NEED(R.num(p.mu_ref) && p.mu_ref >= 0.0, "mu_ref must be >= 0");
  } else if (w == "sutherland") {
    p.mu_model = 1;
    ENTRY("mu_ref");
    NEED(R.num(p.mu_ref) && p.mu_ref > 0.0, "mu_ref must be > 0");
    ENTRY("T_ref");
    NEED(R.num(p.T_ref) && p.T_ref > 0.0, "T_ref must be > 0");
  } else if (w == "power") {
    p.mu_model = 2;
    ENTRY("mu_ref");
    NEED(R.num(p.mu_ref) && p.mu_ref > 0.0, "mu_ref must be > 0");
    ENTRY("T_ref");
    NEED(R.num(p.T_ref) && p.T_ref > 0.0, "T_ref must be > 0");
    ENTRY("omega");
    NEED(R.num(p.omega) && p.omega > 0.0, "omega must be > 0");
  } else { err = "read_material: unknown viscosity type " + w; return false; }
  ENTRY("}");
  ENTRY("prandtl");
  NEED(R.num(p.prandtl) && p.prandtl > 0.0, "prandtl must be > 0");
  ENTRY("flux");
  R.word(w);
  if (w == "kepes_tecno_roe") p.flux = TSM_FLUX_KEPES_TECNO_ROE;
  else if (w == "kepes_tecno_rusanov") p.flux = TSM_FLUX_KEPES_TECNO_RUSANOV;
  else if (w == "kepec") p.flux = TSM_FLUX_KEPEC;
  else if (w == "kep") p.flux = TSM_FLUX_KEP;
  else if (w == "roe_ec") p.flux = TSM_FLUX_ROE_EC;
  else if (w == "simple_avg") p.flux = TSM_FLUX_AVG;
  else { err = "read_material: unknown flux scheme: " + w; return false; }
  BEGIN("flux_par");
  ENTRY("alpha");
  NEED(R.num(p.ev_alpha) && p.ev_alpha >= 0, "alpha must be >= 0");
  ENTRY("beta");
  NEED(R.num(p.ev_beta) && p.ev_beta >= 0, "beta must be >= 0");
  ENTRY("}");
  BEGIN("balance");
  ENTRY("switch_active");
  R.word(w);
  if (w == "yes") {
    NEED(p.model == 1, "read_material: balance switch can be active only with ns model");
    bool b;
    ENTRY("ND"); if (!yesno(R, "ND", b, err)) return false; p.nd_active = b;
    ENTRY("HEAT"); if (!yesno(R, "HEAT", b, err)) return false; p.heat_active = b;
    ENTRY("VISC"); if (!yesno(R, "VISC", b, err)) return false; p.visc_active = b;
  } else if (w == "no") {
    for (int k = 0; k < 6; ++k) R.word(w);  // three "name value" pairs, ignored
    p.visc_active = p.heat_active = 1;
  } else { err = "unknown option " + w + " for switch_active"; return false; }
  ENTRY("}");
  ENTRY("}");
  NEED(!(p.model == 0 && p.mu_ref > 0.0), "read_material:: mu_ref must be 0 for the model=euler:");

  // constants (src/parameter.cc:466-481)
  BEGIN("constants");
  while (!R.eos()) {
    std::string name;
    double v;
    R.word(name);
    NEED(R.num(v), "bad value for constant " + name);
    S.constants[name] = v;
  }

  // initial condition (src/parameter.cc:486-524)
  BEGIN("initial_condition");
  R.word(w);
  if (w == "library") {
    R.word(S.ic_library);
    err = "initial_condition library " + S.ic_library + " is not available (expressions only)";
    return false;
  }
  NEED(w == "temperature", "initial_condition must start with temperature");
  std::string line;
  R.line(line); S.ic["temperature"] = line;
  for (const char* v : {"xvelocity", "yvelocity", "zvelocity", "pressure"}) {
    ENTRY(v);
    R.line(line);
    S.ic[v] = line;
  }
  ENTRY("}");

  // boundary (src/parameter.cc:529-560)
  BEGIN("boundary");
  while (!R.eos()) {
    std::vector<int> tags;
    while (!R.bos()) {
      long t;
      NEED(R.integer(t), "bad boundary tag");
      tags.push_back((int)t);
    }
    ENTRY("type");
    BcSpec b;
    R.word(b.type);
    while (!R.eos()) {
      std::string var;
      R.word(var);
      R.line(line);
      b.fn[var] = line;
    }
    static const char* known[] = {"slip", "noslip", "farfield", "inlet", "outlet", "pressure", "periodic"};
    NEED(std::find(std::begin(known), std::end(known), b.type) != std::end(known),
         "unknown boundary condition " + b.type);
    for (int t : tags) S.bc[t] = b;
  }

  // exact solution (src/parameter.cc:565-607)
  BEGIN("exact_soln");
  ENTRY("available");
  if (!yesno(R, "available", S.exact_available, err)) return false;
  if (S.exact_available) {
    for (const char* v : {"temperature", "xvelocity", "yvelocity", "zvelocity", "pressure"}) {
      ENTRY(v);
      R.line(line);
      S.exact[v] = line;
    }
    ENTRY("}");
  } else {
    while (!R.eos()) R.word(w);
  }

  // integrals (src/parameter.cc:630-653): force blocks, a name and boundary tags
  BEGIN("integrals");
  while (!R.eos()) {
    R.word(w);
    if (w == "force") {
      ENTRY("{");
      std::string name;
      R.word(name);
      std::vector<int> tags;
      while (!R.eos()) {
        long t;
        NEED(R.integer(t), "bad boundary tag in force " + name);
        tags.push_back((int)t);
      }
      NEED(!tags.empty(), "force " + name + " has no boundary tags");
      S.forces.push_back({name, tags});
    }
  }

  // output (src/parameter.cc:642-822)
  BEGIN("output");
  ENTRY("format");
  R.word(S.write_format);
  // vtk (the reference's only format) or raw: little-endian FP64 (T,u,v,p) per
  // vertex behind a 64-byte header, for runs whose ASCII output would dominate
  NEED(S.write_format == "vtk" || S.write_format == "raw", "output format must be vtk or raw");
  ENTRY("frequency");
  long fr;
  NEED(R.integer(fr) && fr > 0, "frequency must be > 0");
  S.write_frequency = (unsigned)fr;
  ENTRY("time_stamps");
  long ts;
  NEED(R.integer(ts) && ts > 0, "time_stamps must be > 0");
  S.n_time_stamps = (int)ts;
  p.n_time_stamps = (int)ts;
  ENTRY("use_online_stat_tool");
  if (!yesno(R, "use_online_stat_tool", S.online_stat, err)) return false;
  BEGIN("variables");
  while (!R.eos()) {
    R.word(w);
    S.write_variables.push_back(w);
  }
  BEGIN("surfaces");
  while (!R.eos()) {
    long id;
    NEED(R.integer(id), "bad surface id");
    S.write_surfaces.push_back((int)id);
  }
  ENTRY("restart");
  if (!yesno(R, "restart", S.write_restart, err)) return false;
  ENTRY("restart_frequency");
  long rf;
  NEED(R.integer(rf), "bad restart_frequency");
  S.restart_frequency = (unsigned)std::max(0L, rf);
  if (S.write_restart) NEED(rf > 0, "restart_frequency must be > 0");
  ENTRY("log_output");
  if (!yesno(R, "log_output", S.write_log, err)) return false;
  ENTRY("soln_output");
  if (!yesno(R, "soln_output", S.write_soln, err)) return false;
  ENTRY("save_global");
  if (!yesno(R, "save_global", S.has_global, err)) return false;
  ENTRY("find_error");
  if (!yesno(R, "find_error", S.find_error, err)) return false;
  ENTRY("}");
  NEED(!(S.sample_list.size() == 1 && S.online_stat),
       "reader_output:: online statistics cannot be evaluated with only one sample.");
  NEED(!(S.write_restart && S.online_stat),
       "reader_output:: restart facility currently unavailable with online statistics.");
  NEED(!(S.find_error && !S.exact_available),
       "reader_output:: Error evaluator activated in the absence of exact solution.");
  return true;
}

}  // namespace tsm
