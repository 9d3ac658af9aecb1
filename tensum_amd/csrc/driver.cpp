// driver.cpp -- `tensum`: drop-in command-line driver over libtsm.
//
// Same command line as the reference (src/process_command_line.cc:23-131):
//   tensum -i param.in -npart 1 -L <dir> [-r] [-p] [-b] [-v] [-d]
// reads <dir>/PARTITION/part_file_0.dat (readme/README.md:155-201), runs every
// sample of the sample list as one batch on the GPU, and writes the
// reference's per-sample outputs in SAMPLE_<id>/: master_file.visit,
// sol<stamp>_0.vtk, residual.dat, global.dat, error_norms.dat,
// restart_<head>_0.dat + LAST_RESTART_HEAD.txt (src/writer.cc:72-409,
// src/fv.cc:1939-2022, :2175-2225, :1712-1795), and with the online
// statistics tool MEAN_ONLINE / VARIANCE_ONLINE / SECOND_MOMENT_ONLINE VTK
// files (src/fv.cc:595-694, :2027-2094).  Errors are printed and give exit
// status 1 (the reference exits with 0, src/mpi_utils.h:29-35).
#include <sys/stat.h>

#include <algorithm>
#include <thread>
#include <mutex>
#include <functional>
#include <deque>
#include <condition_variable>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tsm.h"
#include "expr.h"
#include "param.h"

namespace {

struct PartFile {
  std::vector<double> x, y;
  std::vector<int> btag, bv, cv, pptr, pv;
};

bool die(const std::string& m) {
  std::cerr << "ERROR --> " << m << std::endl;
  return false;
}

// readme/README.md:155-201, src/read_part_mesh_file.cc:43-214 (NPART = 1)
bool read_part_file(const std::string& path, PartFile& P) {
  std::ifstream f(path);
  if (!f) return die("cannot open " + path);
  std::string w;
  int dim, npart, n, ntot;
  f >> w >> dim;
  if (w != "DIMENSION" || dim != 2) return die("part file: DIMENSION 2 expected");
  f >> w >> npart;
  if (w != "PARTITIONS" || npart != 1) return die("part file: one partition expected");
  f >> w >> n >> w >> ntot;
  if (n <= 0 || n != ntot) return die("part file: NODES");
  P.x.resize(n); P.y.resize(n);
  std::vector<int> gid(n);
  std::map<int, int> g2l;
  for (int i = 0; i < n; ++i) {
    double z;
    int ns, part, dbnd;
    f >> gid[i] >> P.x[i] >> P.y[i] >> z >> ns;
    for (int k = 0; k < ns; ++k) f >> part;
    f >> dbnd;
    g2l[gid[i]] = i;
  }
  int np;
  f >> w >> np;
  if (w != "PERIODIC") return die("part file: PERIODIC expected");
  P.pptr.assign(1, 0);
  for (int i = 0; i < np; ++i) {
    int id, cnt, v, nl, part;
    f >> id >> cnt;
    for (int k = 0; k < cnt; ++k) { f >> v; P.pv.push_back(g2l.at(v)); }
    f >> nl;
    for (int k = 0; k < nl; ++k) f >> part;
    P.pptr.push_back((int)P.pv.size());
  }
  int nb, nbt;
  f >> w >> nb >> w >> nbt;
  if (w != "OF") return die("part file: FACES");
  for (int i = 0; i < nb; ++i) {
    int tag, a, b, per;
    f >> tag >> a >> b >> per;
    P.btag.push_back(tag);
    P.bv.push_back(g2l.at(a));
    P.bv.push_back(g2l.at(b));
  }
  int nc, nct;
  f >> w >> nc >> w >> nct;
  for (int i = 0; i < nc; ++i) {
    int a, b, c, ghost;
    f >> a >> b >> c >> ghost;
    if (ghost) return die("part file: ghost cells need NPART > 1");
    P.cv.push_back(g2l.at(a)); P.cv.push_back(g2l.at(b)); P.cv.push_back(g2l.at(c));
  }
  if (!f) return die("part file truncated: " + path);
  return true;
}

bool mkdirs(const std::string& d) {
  if (mkdir(d.c_str(), 0755) == 0 || errno == EEXIST) return true;
  return die("cannot create " + d);
}

// The loop clock a restart file carries after the state (src/writer.cc:404-407).
struct Clock {
  bool valid = false;
  long iter = 0;
  int counter = 0;
  double t = 0.0, res0 = 1.0;
};

// Keeps the lines of a log whose first field is <= n, as the reference's
// awk '$1 <= n' does on a restart (src/fv.cc:350-380): a numeric first field
// compares as a number, any other as a string against n's decimal form; an
// empty line stays.
bool truncate_log(const std::string& path, long n) {
  std::ifstream in(path);
  if (!in) return true;
  std::vector<std::string> keep;
  std::string line;
  const std::string ns = std::to_string(n);
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string f;
    if (!(ls >> f)) { keep.push_back(line); continue; }
    char* end = nullptr;
    const double v = std::strtod(f.c_str(), &end);
    const bool numeric = end && *end == '\0' && end != f.c_str() && std::isfinite(v);
    if (numeric ? v <= (double)n : f <= ns) keep.push_back(line);
  }
  in.close();
  std::ofstream out(path, std::ios::trunc);
  for (const auto& l : keep) out << l << '\n';
  return (bool)out;
}

std::string stamp_name(const std::string& base, int counter) {
  std::ostringstream o;
  o << base << std::setw(4) << std::setfill('0') << counter << "_0";
  return o.str();
}

struct Mat {
  double gamma, R;
  double rho(const double* w) const { return w[3] / (R * w[0]); }
  double mach(const double* w) const { return std::sqrt(w[1] * w[1] + w[2] * w[2]) / std::sqrt(gamma * R * w[0]); }
  double entropy(const double* w) const { return std::log(w[3]) - gamma * std::log(rho(w)); }
};

// Legacy ASCII VTK in the layout of Writer::output_vtk (src/writer.cc:129-328)
using Attached = std::vector<std::pair<std::string, const std::vector<double>*>>;
void write_vtk(const std::string& file, const PartFile& P, const std::vector<double>& W,
               const std::vector<double>& dE, const std::vector<double>& dcarea, const Mat& M,
               const tsm::Settings& S, double time, const Attached& more = {}) {
  std::ofstream vtk(file);
  const int nv = (int)P.x.size(), nc = (int)P.cv.size() / 3;
  vtk << "# vtk DataFile Version 3.0" << std::endl << "tensum" << std::endl << "ASCII" << std::endl
      << "DATASET UNSTRUCTURED_GRID" << std::endl;
  if (S.time_mode == "unsteady")
    vtk << "FIELD FieldData 1" << std::endl << "TIME 1 1 double" << std::endl << time << std::endl;
  vtk << "POINTS  " << nv << "  double" << std::endl;
  for (int i = 0; i < nv; ++i) vtk << P.x[i] << " " << P.y[i] << " " << 0.0 << std::endl;
  vtk << "CELLS  " << nc << " " << 4 * nc << std::endl;
  for (int i = 0; i < nc; ++i)
    vtk << 3 << " " << P.cv[3 * i] << " " << P.cv[3 * i + 1] << " " << P.cv[3 * i + 2] << std::endl;
  vtk << "CELL_TYPES  " << nc << std::endl;
  for (int i = 0; i < nc; ++i) vtk << 5 << std::endl;
  vtk << "POINT_DATA  " << nv << std::endl;
  vtk << "SCALARS pressure double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
  for (int i = 0; i < nv; ++i) vtk << std::setprecision(12) << W[4 * i + 3] << std::endl;
  vtk << "SCALARS temperature double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
  for (int i = 0; i < nv; ++i) vtk << std::setprecision(12) << W[4 * i] << std::endl;
  vtk << "VECTORS velocity double" << std::endl;
  for (int i = 0; i < nv; ++i)
    vtk << std::setprecision(12) << W[4 * i + 1] << " " << W[4 * i + 2] << " " << 0.0 << std::endl;
  auto has = [&](const char* n) {
    for (auto& v : S.write_variables) if (v == n) return true;
    return false;
  };
  if (has("mach")) {
    vtk << "SCALARS mach double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
    for (int i = 0; i < nv; ++i) vtk << std::setprecision(12) << M.mach(&W[4 * i]) << std::endl;
  }
  if (has("density")) {
    vtk << "SCALARS density double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
    for (int i = 0; i < nv; ++i) vtk << std::setprecision(12) << M.rho(&W[4 * i]) << std::endl;
  }
  if (has("entropy")) {
    vtk << "SCALARS entropy double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
    for (int i = 0; i < nv; ++i) vtk << std::setprecision(12) << M.entropy(&W[4 * i]) << std::endl;
  }
  if (has("vorticity")) {  // src/material.h:622-627 in (f, vx, vy, l) x (x, y) layout
    vtk << "SCALARS vorticity double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
    for (int i = 0; i < nv; ++i) {
      const double* g = &dE[8 * i];
      double om = M.R * W[4 * i] * (g[4] + W[4 * i + 2] * g[6] - g[3] - W[4 * i + 1] * g[7]);
      vtk << std::setprecision(12) << om << std::endl;
    }
  }
  vtk << "SCALARS  dual_area  double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
  for (int i = 0; i < nv; ++i) vtk << std::setprecision(12) << dcarea[i] << std::endl;
  for (auto& a : more) {  // further attached vertex data, in attach order
    vtk << "SCALARS  " << a.first << "  double 1" << std::endl << "LOOKUP_TABLE default" << std::endl;
    for (int i = 0; i < nv; ++i) vtk << std::setprecision(12) << (*a.second)[i] << std::endl;
  }
}

// Vorticity from the entropy-variable gradients (src/material.h:622-627) in the
// (f, vx, vy, l) x (x, y) layout.
double vorticity(const Mat& M, const double* w, const double* g) {
  return M.R * w[0] * (g[4] + w[2] * g[6] - g[3] - w[1] * g[7]);
}

// Raw solution file (format raw): "TSMRAW01", int64 n_vertex, int64 n_fields
// (4), double time, 32 bytes of field names "T u v p", then n_vertex x 4
// little-endian doubles in part-file vertex order.
bool write_raw(const std::string& file, const std::vector<double>& W, int nv, double time) {
  std::ofstream f(file, std::ios::binary);
  if (!f) return false;
  char head[64] = {0};
  std::memcpy(head, "TSMRAW01", 8);
  const long long n = nv, nf = 4;
  std::memcpy(head + 8, &n, 8);
  std::memcpy(head + 16, &nf, 8);
  std::memcpy(head + 24, &time, 8);
  std::memcpy(head + 32, "T u v p", 7);
  f.write(head, 64);
  f.write(reinterpret_cast<const char*>(W.data()), (std::streamsize)(4 * (size_t)nv * sizeof(double)));
  return (bool)f;
}

// Files are written on a thread of their own, in the order they are posted,
// while the GPU carries on; a bounded queue keeps the copies held in memory
// in check.
class AsyncWriter {
 public:
  AsyncWriter() : th_([this] { loop(); }) {}
  ~AsyncWriter() { finish(); }
  void post(std::function<bool()> job) {
    std::unique_lock<std::mutex> l(m_);
    space_.wait(l, [this] { return q_.size() < 16; });
    q_.push_back(std::move(job));
    ready_.notify_one();
  }
  bool finish() {  // drains the queue; false if any write failed
    {
      std::lock_guard<std::mutex> l(m_);
      done_ = true;
      ready_.notify_one();
    }
    if (th_.joinable()) th_.join();
    return ok_;
  }

 private:
  void loop() {
    for (;;) {
      std::function<bool()> job;
      {
        std::unique_lock<std::mutex> l(m_);
        ready_.wait(l, [this] { return done_ || !q_.empty(); });
        if (q_.empty()) return;
        job = std::move(q_.front());
        q_.pop_front();
        space_.notify_one();
      }
      if (!job()) ok_ = false;
    }
  }
  std::mutex m_;
  std::condition_variable ready_, space_;
  std::deque<std::function<bool()>> q_;
  bool done_ = false;
  std::atomic<bool> ok_{true};
  std::thread th_;
};

// Writer threads for many samples: a sample's files always go to the same
// thread, so each sample's files (and master file lines) keep their order.
class WriterPool {
 public:
  explicit WriterPool(int n) : w_(std::max(1, n)) {}
  void post(int sample, std::function<bool()> job) { w_[sample % w_.size()].post(std::move(job)); }
  bool finish() {
    bool ok = true;
    for (auto& w : w_) ok = w.finish() && ok;
    return ok;
  }

 private:
  std::deque<AsyncWriter> w_;
};

struct Moments {  // src/fv.cc:595-694 (primitive fields + density/mach/entropy)
  std::vector<std::vector<double>> fm, sm, var;  // per stamp: nv * 7
};

}  // namespace

int main(int argc, char** argv) {
  std::string ifile, loc;
  int npart = -1;
  bool restart = false, check_only = false, bounds = false, enstrophy = false;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-i" && i + 1 < argc) ifile = argv[++i];
    else if (a == "-npart" && i + 1 < argc) npart = std::atoi(argv[++i]);
    else if (a == "-L" && i + 1 < argc) loc = argv[++i];
    else if (a == "-r") restart = true;
    else if (a == "--check") { check_only = true; npart = npart < 1 ? 1 : npart; loc = loc.empty() ? "." : loc; }
    else if (a == "-b") bounds = true;
    else if (a == "-e") enstrophy = true;
    else if (a == "-p" || a == "-v" || a == "-d") {}
    else if (a == "-axi") { std::cerr << "ERROR --> -axi is not supported (2-D only)" << std::endl; return 1; }
    else { std::cerr << "Unknown command line flag: " << a << std::endl; ifile.clear(); break; }
  }
  if (ifile.empty() || loc.empty() || npart < 1) {
    std::cout << "Valid flags are:\n   -i filename   Specify input file name (required)\n"
                 "   -npart N      Specify number of mesh partitions N (required, 1)\n"
                 "   -L path       Location of mesh partition folder (required)\n"
                 "   -r            Restart from SAMPLE_<id>/restart files\n"
                 "   -b            Write the solution bounds to SAMPLE_<id>/bounds.dat\n"
                 "   -e            With save_global: enstrophy as a fifth column of global.dat\n";
    return 1;
  }
  if (npart != 1) { std::cerr << "ERROR --> one GPU holds the whole mesh: use -npart 1" << std::endl; return 1; }
  tsm::Settings S;
  std::string err;
  if (!tsm::read_settings(ifile, S, err)) { std::cerr << "ERROR --> " << err << std::endl; return 1; }
  if (check_only) {  // parse report, no device use
    const tsm_params& p = S.p;
    std::cout << std::setprecision(17) << "cell_type " << p.cell_type << "\ntime_mode " << S.time_mode
              << "\ntime_scheme " << S.time_scheme << "\ncfl " << p.cfl << "\ntime_step " << p.time_step
              << "\nfinal_time " << p.final_time << "\nmax_iter " << p.max_iter << "\nreconstruct "
              << p.reconstruct << "\ntvb_M " << p.tvb_M << "\nliou_fix " << p.liou_fix << "\ngamma "
              << p.gamma << "\ngas_const " << p.gas_const << "\nmodel " << p.model << "\nmu_model "
              << p.mu_model << "\nmu_ref " << p.mu_ref << "\nprandtl " << p.prandtl << "\nflux " << p.flux
              << "\nev_alpha " << p.ev_alpha << "\nev_beta " << p.ev_beta << "\nnd_active " << p.nd_active
              << "\nn_samples " << S.sample_list.size() << "\nrnd_per_sample " << S.rnd_per_sample
              << "\ntime_stamps " << S.n_time_stamps << "\nonline_stat " << S.online_stat
              << "\nwrite_restart " << S.write_restart << "\nhas_global " << S.has_global
              << "\nfind_error " << S.find_error << "\nn_constants " << S.constants.size()
              << "\nn_bc " << S.bc.size() << "\n";
    for (auto& kv : S.bc) std::cout << "bc " << kv.first << " " << kv.second.type << "\n";
    for (auto& kv : S.constants) std::cout << "const " << kv.first << " " << kv.second << "\n";
    for (const char* v : {"temperature", "xvelocity", "yvelocity", "pressure"}) {
      tsm::Expr e;
      if (!e.parse(S.ic[v], S.constants, err)) { std::cerr << "ERROR --> " << err << std::endl; return 1; }
      e.set_pert([](int, double, double, double k) { return 0.1 * k; });
      std::cout << "ic " << v << " " << e.eval(0.25, -0.5, 0.0) << "\n";
    }
    return 0;
  }
  PartFile P;
  if (!read_part_file(loc + "/PARTITION/part_file_0.dat", P)) return 1;
  const int nv = (int)P.x.size();

  // boundary table: constant boundary states from the expressions
  std::vector<int> bc_tag, bc_type, bc_adiabatic;
  std::vector<double> bc_state;
  const std::map<std::string, int> bcode = {{"slip", 1}, {"noslip", 2}, {"farfield", 3}, {"inlet", 4},
                                            {"outlet", 5}, {"pressure", 6}, {"periodic", 7}};
  // Each boundary's expressions (src/bc.h:411-562) are evaluated at every
  // bface end point with each sample's random numbers once the samples are
  // known (below); the table holds the values at the tag's first end point
  // for sample 0.  Expressions of t are handed to the library as text, which
  // re-evaluates them on the device at every stage time (Heun only).
  const char* bnames[4] = {"temperature", "xvelocity", "yvelocity", "pressure"};
  std::map<int, std::vector<tsm::Expr>> bexpr;   // tag -> 4 expressions (empty: field not given)
  std::map<int, std::vector<bool>> bhas;
  std::map<int, std::vector<std::string>> btime; // tag -> texts of its fields of t ("" none)
  for (auto& kv : S.bc) {
    bc_tag.push_back(kv.first);
    bc_type.push_back(bcode.at(kv.second.type));
    std::vector<tsm::Expr>& ex = bexpr[kv.first];
    std::vector<bool>& has = bhas[kv.first];
    ex.resize(4);
    has.assign(4, false);
    for (int k = 0; k < 4; ++k) {
      auto it = kv.second.fn.find(bnames[k]);
      if (it == kv.second.fn.end()) continue;
      if (!ex[k].parse(it->second, S.constants, err)) { std::cerr << "ERROR --> " << err << std::endl; return 1; }
      if (ex[k].uses_time()) {
        btime[kv.first].resize(4);
        btime[kv.first][k] = it->second;
      }
      has[k] = true;
    }
    double st[4] = {0, 0, 0, 0};
    bc_state.insert(bc_state.end(), st, st + 4);
    // a noslip wall without a temperature is adiabatic (src/bc.h:246-276)
    bc_adiabatic.push_back(kv.second.type == "slip" ||
                           (kv.second.type == "noslip" && !kv.second.fn.count("temperature")));
  }
  tsm_mesh m{};
  m.n_vertex = nv; m.x = P.x.data(); m.y = P.y.data();
  m.n_bface = (int)P.btag.size(); m.bface_tag = P.btag.data(); m.bface_v = P.bv.data();
  m.n_cell = (int)P.cv.size() / 3; m.cell_v = P.cv.data();
  m.n_periodic = (int)P.pptr.size() - 1; m.periodic_ptr = P.pptr.data(); m.periodic_v = P.pv.data();
  m.n_bc = (int)bc_tag.size(); m.bc_tag = bc_tag.data(); m.bc_type = bc_type.data(); m.bc_state = bc_state.data();
  m.bc_adiabatic = bc_adiabatic.data();

  // samples and their random inputs (src/fv.cc:502-551)
  const int ns = (int)S.sample_list.size();
  std::vector<std::vector<double>> rnd(ns);
  if (S.rnd_per_sample > 0) {
    std::ifstream rf(S.rnd_file_loc);
    if (!rf) { std::cerr << "ERROR --> the random parameter file " << S.rnd_file_loc << " doesn't exist!!" << std::endl; return 1; }
    int total;
    rf >> total;
    std::map<int, std::vector<double>> rows;
    for (int r = 0; r < total; ++r) {
      int id;
      rf >> id;
      std::vector<double> v(S.rnd_per_sample);
      for (auto& q : v) rf >> q;
      rows[id] = v;
    }
    for (int k = 0; k < ns; ++k) {
      auto it = rows.find(S.sample_list[k]);
      if (it == rows.end()) { std::cerr << "ERROR --> Sample " << S.sample_list[k] << " not present in " << S.rnd_file_loc << std::endl; return 1; }
      rnd[k] = it->second;
    }
  }
  // boundary values per sample and bface end point
  const int nbf = (int)P.btag.size();
  std::vector<std::vector<double>> bvals(ns, std::vector<double>(8 * (size_t)nbf, 0.0));
  for (int k = 0; k < ns; ++k) {
    const std::vector<double>& r = rnd[k];
    auto pert = [&r](int type, double px, double py, double kk) {
      double v = 0.0;  // src/perturb.cc:16-209
      if (tsm_pert(type, px, py, kk, r.data(), (int)r.size(), &v) != TSM_OK) return (double)NAN;
      return v;
    };
    for (int b = 0; b < nbf; ++b) {
      auto& ex = bexpr[P.btag[b]];
      const auto& has = bhas[P.btag[b]];
      for (int e = 0; e < 2; ++e) {
        const int v = P.bv[2 * b + e];
        for (int q = 0; q < 4; ++q) {
          if (!has[q]) continue;
          ex[q].set_pert(pert);
          bvals[k][8 * (size_t)b + 4 * e + q] = ex[q].eval(P.x[v], P.y[v], 0.0);
        }
      }
    }
  }
  // table: each tag's values at its first end point for sample 0; per-sample
  // values are passed only if some value differs from its tag's table entry
  std::vector<int> btab(nbf);
  std::vector<bool> filled(bc_tag.size(), false);
  for (int b = 0; b < nbf; ++b) {
    btab[b] = (int)(std::find(bc_tag.begin(), bc_tag.end(), P.btag[b]) - bc_tag.begin());
    if (ns > 0 && !filled[btab[b]]) {
      for (int q = 0; q < 4; ++q) bc_state[4 * (size_t)btab[b] + q] = bvals[0][8 * (size_t)b + q];
      filled[btab[b]] = true;
    }
  }
  bool varying = false;
  for (int k = 0; k < ns && !varying; ++k)
    for (int b = 0; b < nbf && !varying; ++b)
      for (int i = 0; i < 8; ++i)
        if (bvals[k][8 * (size_t)b + i] != bc_state[4 * (size_t)btab[b] + i % 4]) varying = true;
  tsm_ctx* ctx = nullptr;
  if (tsm_create(&m, &S.p, ns, &ctx) != TSM_OK) {
    std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
    return 1;
  }
  if (varying)
    for (int k = 0; k < ns; ++k)
      if (tsm_set_bc_state(ctx, k, bvals[k].data()) != TSM_OK) {
        std::cerr << "ERROR --> boundary values of SAMPLE_" << S.sample_list[k] << ": " << tsm_last_error() << std::endl;
        return 1;
      }
  if (!btime.empty()) {  // boundary values of t: one group per tag that has them
    std::vector<int> group(nbf, -1);
    std::vector<const char*> texts;
    std::vector<int> gtag;
    for (auto& kv : btime) {
      gtag.push_back(kv.first);
      for (int q = 0; q < 4; ++q) texts.push_back(kv.second[q].empty() ? nullptr : kv.second[q].c_str());
    }
    for (int b = 0; b < nbf; ++b)
      group[b] = (int)(std::find(gtag.begin(), gtag.end(), P.btag[b]) - gtag.begin());
    for (int& g : group)
      if (g == (int)gtag.size()) g = -1;
    std::vector<const char*> cname;
    std::vector<double> cval;
    for (auto& kv : S.constants) { cname.push_back(kv.first.c_str()); cval.push_back(kv.second); }
    if (tsm_set_bc_time_expr(ctx, group.data(), (int)gtag.size(), texts.data(), (int)cname.size(),
                             cname.data(), cval.data()) != TSM_OK) {
      std::cerr << "ERROR --> boundary values of t: " << tsm_last_error() << std::endl;
      return 1;
    }
  }
  const Mat M{S.p.gamma, S.p.gas_const};
  std::vector<double> dcarea(nv), mcarea(nv);
  tsm_geometry(ctx, nullptr, nullptr, dcarea.data(), mcarea.data());
  std::vector<std::string> dirs(ns);
  std::vector<std::ofstream> res_f(ns), glob_f(ns), force_f(ns), bnd_f(ns);
  // every force's tags must name boundary faces of the mesh (src/force.cc:25-106)
  for (auto& fo : S.forces)
    for (int t : fo.second)
      if (std::find(P.btag.begin(), P.btag.end(), t) == P.btag.end()) {
        std::cerr << "Force: " << fo.first << " has face type " << t << " but it is not present in the grid" << std::endl;
        return 1;
      }
  std::vector<int> head(ns, 0);
  std::vector<double> W(4 * (size_t)nv), dE(8 * (size_t)nv);
  const bool steady = S.time_mode == "steady";
  std::vector<Clock> rc_clock(ns);
  for (int k = 0; k < ns; ++k) {
    dirs[k] = "SAMPLE_" + std::to_string(S.sample_list[k]);
    // a run without -r starts from an empty sample directory (src/fv.cc:226-231)
    if (!restart) {
      std::error_code ec;
      std::filesystem::remove_all(dirs[k], ec);
    }
    if (!mkdirs(dirs[k])) return 1;
    // initial state: restart file or expressions (src/fv.cc:300-402)
    bool from_restart = false;
    if (restart) {
      std::ifstream hf(dirs[k] + "/LAST_RESTART_HEAD.txt");
      int h;
      if (hf >> h) {
        std::ifstream rf(dirs[k] + "/restart_" + std::to_string(h) + "_0.dat");
        double wz;
        for (int i = 0; i < nv; ++i) rf >> W[4 * i] >> W[4 * i + 1] >> W[4 * i + 2] >> wz >> W[4 * i + 3];
        rc_clock[k].valid = (bool)(rf >> rc_clock[k].iter >> rc_clock[k].counter >> rc_clock[k].t >> rc_clock[k].res0);
        if (!rc_clock[k].valid) { std::cerr << "ERROR --> bad restart file in " << dirs[k] << std::endl; return 1; }
        from_restart = true;
        const Clock& q = rc_clock[k];
        if (q.t >= S.p.final_time)
          std::cout << "    SAMPLE " << S.sample_list[k] << " has already reached final time!!" << std::endl;
        else if (q.iter >= S.p.max_iter && steady)
          std::cout << "    SAMPLE " << S.sample_list[k] << " has already max iteration count!!" << std::endl;
        else if (q.t > 0.0) {
          // the logs keep the lines of the steps up to the restart (src/fv.cc:350-380)
          if (S.write_log) truncate_log(dirs[k] + "/residual.dat", q.iter);
          if (!S.forces.empty()) truncate_log(dirs[k] + "/force.dat", q.iter);
          if (S.has_global) truncate_log(dirs[k] + "/global.dat", q.iter);
        }
      } else {
        std::cout << "    Restart file(s) for " << dirs[k] << " does not exist. Staring from T=0 ..." << std::endl;
      }
    }
    if (!from_restart) {
      tsm::Expr e[4];
      const char* names[4] = {"temperature", "xvelocity", "yvelocity", "pressure"};
      const std::vector<double>& r = rnd[k];
      for (int q = 0; q < 4; ++q) {
        if (!e[q].parse(S.ic[names[q]], S.constants, err)) { std::cerr << "ERROR --> " << err << std::endl; return 1; }
        e[q].set_pert([&r](int type, double px, double py, double kk) {
          double v = 0.0;  // src/perturb.cc:16-209
          if (tsm_pert(type, px, py, kk, r.data(), (int)r.size(), &v) != TSM_OK) return (double)NAN;
          return v;
        });
      }
      for (int i = 0; i < nv; ++i)
        for (int q = 0; q < 4; ++q) W[4 * i + q] = e[q].eval(P.x[i], P.y[i], 0.0);
      for (int i = 0; i < 4 * nv; ++i)
        if (!std::isfinite(W[i])) {
          std::cerr << "ERROR --> initial condition of " << dirs[k]
                    << " is not finite (unknown pert type or too few random numbers?)" << std::endl;
          return 1;
        }
    }
    if (tsm_set_state(ctx, k, W.data()) != TSM_OK) { std::cerr << "ERROR --> " << tsm_last_error() << std::endl; return 1; }
    if (rc_clock[k].valid &&
        tsm_set_clock(ctx, k, rc_clock[k].iter, rc_clock[k].counter, rc_clock[k].t, rc_clock[k].res0) != TSM_OK) {
      std::cerr << "ERROR --> restart clock of " << dirs[k] << ": " << tsm_last_error() << std::endl;
      return 1;
    }
    if (!restart) std::ofstream(dirs[k] + "/master_file.visit") << "!NBLOCKS 1" << std::endl;
    if (S.write_log) res_f[k].open(dirs[k] + "/residual.dat", std::ios::app);
    if (bounds) {  // -b: solution bounds after every logged step (src/fv.cc:258-278)
      bnd_f[k].open(dirs[k] + "/bounds.dat", std::ios::app);
      if (!restart) {
        bnd_f[k] << std::left << std::setw(16 * 11) << std::setfill('-') << "" << std::endl;
        for (const char* h : {"TIME", "TEMP_MIN", "TEMP_MAX", "P_MIN", "P_MAX", "U_MIN", "U_MAX", "V_MIN",
                              "V_MAX", "W_MIN", "W_MAX"})
          bnd_f[k] << std::left << std::setw(17) << std::setfill(' ') << h;
        bnd_f[k] << std::endl << std::left << std::setw(16 * 11) << std::setfill('-') << "" << std::endl;
      }
    }
    if (!S.forces.empty()) force_f[k].open(dirs[k] + "/force.dat", std::ios::app);
    if (S.has_global) glob_f[k].open(dirs[k] + "/global.dat", std::ios::app);
  }
  // a steady run lists only its final solution in the master file and keeps
  // overwriting sol0001 every write_frequency steps (src/fv.cc:2241-2245, :2318-2323, :2371-2377)
  WriterPool writer(std::min(ns, std::max(1, std::min(8, (int)std::thread::hardware_concurrency() - 1))));
  const bool raw = S.write_format == "raw";
  auto listed = [&](const char* n) {
    return std::find(S.write_variables.begin(), S.write_variables.end(), n) != S.write_variables.end();
  };
  // The outputs after a step read the gradient array as the loop body left it:
  // the gradients of the last stage's input state (src/fv.cc:1476, :2295-2304),
  // or of the current state where a Navier-Stokes force evaluation recomputed
  // them (src/force.cc:124).  Before any step both are those of the initial state.
  const bool ns_model = S.p.model == 1;
  const bool fresh_grad = ns_model && !S.forces.empty();
  const bool surf_ns = ns_model && !S.write_surfaces.empty();
  const bool kept = !fresh_grad && (listed("vorticity") || surf_ns);
  if (kept && tsm_keep_stage_state(ctx, 1) != TSM_OK) {
    std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
    return 1;
  }
  auto out_gradients = [&](int k, double* g) {
    return kept ? tsm_stage_gradients(ctx, k, g) : tsm_gradients(ctx, k, g);
  };
  // surfaces (src/writer.cc:336-377, src/output_skin_friction.cc, src/output_heatflux.cc):
  // the reference's bface order is the part file's, its vertex order the
  // preprocessed one
  std::vector<int> bfv;
  if (!S.write_surfaces.empty()) {
    int nf = 0, nbq = 0;
    bfv.assign(2 * (size_t)nbf, 0);
    if (tsm_host_geometry(&m, S.p.cell_type, nullptr, nullptr, bfv.data(), nullptr, nullptr, nullptr,
                          &nf, &nbq) != TSM_OK || nbq != nbf) {
      std::cerr << "ERROR --> surfaces: " << tsm_last_error() << std::endl;
      return 1;
    }
  }
  auto surfaces = [&](int k, int counter) {
    if (S.write_surfaces.empty()) return;
    const std::string idx = stamp_name("", counter);
    std::vector<std::vector<double>> sq(S.write_surfaces.size());
    for (size_t i = 0; i < S.write_surfaces.size() && ns_model; ++i) {
      const int tag = S.write_surfaces[i];
      sq[i].assign(2 * (size_t)std::count(P.btag.begin(), P.btag.end(), tag), 0.0);
      if (tsm_surface(ctx, k, kept ? 1 : 0, &tag, 1, sq[i].data()) != TSM_OK)
        std::cerr << "ERROR --> surfaces: " << tsm_last_error() << std::endl;
    }
    writer.post(k, [&P, &S, &bfv, ns_model, idx, dir = dirs[k], w = W, sq]() {
      bool good = true;
      for (size_t i = 0; i < S.write_surfaces.size(); ++i) {
        const int tag = S.write_surfaces[i];
        const std::string sfx = idx + "_" + std::to_string(tag) + ".dat";
        std::ofstream fv(dir + "/v" + sfx);
        for (size_t b = 0; b < P.btag.size(); ++b) {
          if (P.btag[b] != tag) continue;
          const int v = bfv[2 * b];
          const double* q = &w[4 * (size_t)v];
          fv << P.x[v] << "  " << P.y[v] << "  " << q[3] << "  " << q[0] << "  " << q[1] << "  " << q[2]
             << "  " << 0.0 << "  " << std::endl;
        }
        good = good && (bool)fv;
        if (!ns_model) continue;
        std::ofstream ff(dir + "/f" + sfx), fq(dir + "/q" + sfx);
        size_t j = 0;
        for (size_t b = 0; b < P.btag.size(); ++b) {
          if (P.btag[b] != tag) continue;
          const int v0 = bfv[2 * b], v1 = bfv[2 * b + 1];
          const double cx = P.x[v0] * 0.5 + P.x[v1] * 0.5, cy = P.y[v0] * 0.5 + P.y[v1] * 0.5;
          ff << cx << "  " << cy << "  " << sq[i][2 * j] << "  " << std::endl;
          fq << cx << "  " << cy << "  " << sq[i][2 * j + 1] << "  " << std::endl;
          ++j;
        }
        good = good && (bool)ff && (bool)fq;
      }
      return good;
    });
  };
  // a failed download ends the run (no stale state is written): false after
  // printing the library's message
  auto output = [&](int k, int counter, double t, bool master = true) {
    if (tsm_get_state(ctx, k, W.data()) != TSM_OK || (!raw && out_gradients(k, dE.data()) != TSM_OK)) {
      std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
      return false;
    }
    const std::string name = stamp_name("sol", counter) + (raw ? ".raw" : ".vtk");
    const std::string dir = dirs[k];
    writer.post(k, [&P, &dcarea, &M, &S, raw, nv, dir, name, t, master, w = W,
                 g = raw ? std::vector<double>() : dE]() {
      const bool good = raw ? write_raw(dir + "/" + name, w, nv, t)
                            : (write_vtk(dir + "/" + name, P, w, g, dcarea, M, S, t), true);
      if (master) std::ofstream(dir + "/master_file.visit", std::ios::app) << name << std::endl;
      return good;
    });
    surfaces(k, counter);
    return true;
  };
  // online statistics across samples at each stamp (Welford in sample order);
  // a steady run has two: the initial and the converged state
  const int nst = steady ? 2 : S.n_time_stamps + 1;
  std::vector<std::vector<std::vector<double>>> snap;
  if (S.online_stat) snap.assign(nst, std::vector<std::vector<double>>(ns));
  const bool has_vort = listed("vorticity");
  std::vector<std::vector<std::vector<double>>> vsnap;  // per stamp and sample: vorticity
  if (S.online_stat && has_vort) vsnap.assign(nst, std::vector<std::vector<double>>(ns));
  auto keep = [&](int k, int counter) {
    if (!S.online_stat) return true;
    if (tsm_get_state(ctx, k, W.data()) != TSM_OK ||
        (has_vort && out_gradients(k, dE.data()) != TSM_OK)) {
      std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
      return false;
    }
    snap[counter][k] = W;
    if (has_vort) {
      std::vector<double> om(nv);
      for (int i = 0; i < nv; ++i) om[i] = vorticity(M, &W[4 * i], &dE[8 * i]);
      vsnap[counter][k] = std::move(om);
    }
    return true;
  };
  for (int k = 0; k < ns; ++k) {
    if (S.write_soln && !restart && !output(k, 0, 0.0)) return 1;  // src/fv.cc:2240-2245
    if (!keep(k, 0)) return 1;
  }
  // time loop, all samples together; outputs as each reaches its stamps
  std::vector<double> logbuf(4 * (size_t)ns);
  std::vector<long> done(ns);
  std::vector<int> cnt_before(ns, 1), finished(ns, 0);
  std::vector<double> st_v(3 * (size_t)ns), st_rn(4 * (size_t)ns), g_ke(ns), g_ent(ns);
  std::vector<long> st_iter(ns);
  std::vector<int> st_cnt(ns), st_done(ns);
  for (int k = 0; k < ns; ++k)
    if (rc_clock[k].valid) cnt_before[k] = std::max(1, rc_clock[k].counter);
  int live = ns;
  while (live > 0) {
    if (tsm_run(ctx, 1, done.data(), logbuf.data()) != TSM_OK) {
      std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
      return 1;
    }
    // every sample's loop state, and its integrals, once per step
    if (tsm_status_all(ctx, st_v.data(), st_iter.data(), st_cnt.data(), st_done.data(), st_rn.data()) != TSM_OK ||
        (S.has_global && tsm_global_all(ctx, g_ke.data(), g_ent.data()) != TSM_OK)) {
      std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
      return 1;
    }
    for (int k = 0; k < ns; ++k) {
      if (finished[k]) continue;
      const double* v = &st_v[3 * (size_t)k];
      const long iter = st_iter[k];
      const int counter = st_cnt[k], dn = st_done[k];
      if (done[k] > 0) {
        const double* L = &logbuf[4 * (size_t)k];
        if (S.write_log && steady) {  // src/fv.cc:1944-1957
          const double* rn = &st_rn[4 * (size_t)k];
          res_f[k] << std::setw(10) << (long)L[0] << "  " << std::scientific << std::setprecision(4)
                   << L[1] << "  " << L[3] << "  " << rn[0] << "  " << rn[1] << "  " << rn[2] << "  "
                   << 0.0 << "  " << rn[3] << std::endl;
        } else if (S.write_log) {
          res_f[k] << std::setw(10) << (long)L[0] << "  " << std::scientific << std::setprecision(12)
                   << L[1] << "  " << L[2] << "  " << L[3] << std::endl;
        }
        if (bounds && S.write_log) {  // src/fv.cc:1971-1972, :2099-2171
          double b[10];
          if (tsm_bounds(ctx, k, b) != TSM_OK) {
            std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
            return 1;
          }
          bnd_f[k] << std::left << std::setw(17) << std::setfill(' ') << std::setprecision(10) << L[2];
          for (double x : b) bnd_f[k] << std::left << std::setw(17) << std::setfill(' ') << std::setprecision(10) << x;
          bnd_f[k] << std::endl;
          if (b[0] < 0.0 || b[2] < 0.0) {
            if (S.write_soln) (void)output(k, counter, L[2]);
            writer.finish();
            std::cerr << "ERROR --> Error:: Observed negative temperature/pressure in solution (" << dirs[k]
                      << ")" << std::endl;
            return 1;
          }
        }
        if (!S.forces.empty()) {  // src/force.cc:118-217, after each step
          std::ofstream& ff = force_f[k];
          ff << std::setw(6) << (long)L[0] << " " << std::scientific << std::setw(15);
          ff << L[2] << std::setw(15);
          for (size_t i = 0; i < S.forces.size(); ++i) {
            double fo[4];
            const auto& tags = S.forces[i].second;
            if (tsm_forces(ctx, k, (int)i, tags.data(), (int)tags.size(), fo) != TSM_OK) {
              std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
              return 1;
            }
            ff << fo[0] << "  " << fo[1] << "  " << fo[2] << "  " << fo[3] << std::endl;
          }
        }
        if (S.has_global) {
          glob_f[k] << std::defaultfloat << std::setprecision(6) << (long)L[0] << "  " << L[2]
                    << "  " << g_ke[k] << "  " << g_ent[k];
          if (enstrophy) {  // no counterpart in the reference's global.dat
            double z = 0.0;
            if (tsm_enstrophy(ctx, k, &z) != TSM_OK) {
              std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
              return 1;
            }
            glob_f[k] << "  " << z;
          }
          glob_f[k] << std::endl;
        }
        if (steady) {
          if (S.write_soln && iter % S.write_frequency == 0 && !output(k, 1, v[0], false)) return 1;
        } else if (counter > cnt_before[k]) {
          if (S.write_soln && !output(k, cnt_before[k], v[0])) return 1;
          if (!keep(k, cnt_before[k])) return 1;
          cnt_before[k] = counter;
        }
        if (S.write_restart && (iter % S.restart_frequency == 0 || v[0] == S.p.final_time)) {
          double res0 = 1.0;
          if (tsm_get_state(ctx, k, W.data()) != TSM_OK ||
              tsm_get_clock(ctx, k, nullptr, nullptr, nullptr, &res0) != TSM_OK) {
            std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
            return 1;
          }
          writer.post(k, [nv, dir = dirs[k], h = head[k], iter, counter, t = v[0], res0, w = W]() {
            std::ofstream fo(dir + "/restart_" + std::to_string(h) + "_0.dat");
            for (int i = 0; i < nv; ++i)
              fo << std::scientific << std::setprecision(15) << w[4 * i] << "  " << w[4 * i + 1] << "  "
                 << w[4 * i + 2] << "  " << 0.0 << "  " << w[4 * i + 3] << std::endl;
            fo << iter << std::endl << counter << std::endl << t << std::endl << res0 << std::endl;
            const bool good = (bool)fo;
            fo.close();
            std::ofstream(dir + "/LAST_RESTART_HEAD.txt") << h << std::endl;
            return good;
          });
          head[k] = (head[k] + 1) % 2;
        }
      }
      if (dn || done[k] == 0) {
        finished[k] = 1;
        --live;
        if (steady) {  // final output of a steady run (src/fv.cc:2371-2382)
          if (S.write_soln && !output(k, 1, v[0])) return 1;
          if (!keep(k, 1)) return 1;
        }
        if (!std::isfinite(v[2]))
          std::cerr << "    " << dirs[k] << ": non-finite residual at t = " << v[0] << ", time loop left" << std::endl;
      }
    }
  }
  // error norms (src/fv.cc:1712-1795)
  if (S.find_error) {
    for (int k = 0; k < ns; ++k) {
      tsm::Expr e[4];
      const char* names[4] = {"temperature", "xvelocity", "yvelocity", "pressure"};
      for (int q = 0; q < 4; ++q)
        if (!e[q].parse(S.exact[names[q]], S.constants, err)) { std::cerr << "ERROR --> " << err << std::endl; return 1; }
      std::vector<double> ex(4 * (size_t)nv);
      for (int i = 0; i < nv; ++i)
        for (int q = 0; q < 4; ++q) ex[4 * i + q] = e[q].eval(P.x[i], P.y[i], S.p.final_time);
      double o[24];
      if (tsm_error_norms(ctx, k, ex.data(), o) != TSM_OK) {
        std::cerr << "ERROR --> " << tsm_last_error() << std::endl;
        return 1;
      }
      std::ofstream ef(dirs[k] + "/error_norms.dat");
      const char* lab[8] = {"TEMPERATURE", "PRESSURE", "VEL_X", "VEL_Y", "VEL_Z", "DENSITY", "MACH", "ENTROPY"};
      ef << "N = " << nv << "\n" << std::scientific << std::setprecision(8);
      for (int q = 0; q < 8; ++q)
        ef << lab[q] << " (L1, L2, Linf): " << o[3 * q] << "   " << o[3 * q + 1] << "   " << o[3 * q + 2] << "\n";
    }
  }
  // moments across samples, in sample order (src/fv.cc:595-694, src/reduce_moments.cc:115-141):
  // the primitive state, and density, Mach number, vorticity and entropy when
  // listed in `variables`; their files (src/fv.cc:2027-2094) with master files,
  // and the L1 norms of the mean and of the second moment
  // (src/fv.cc:1884-1934, src/save_stat_norm.cc:14-100)
  if (S.online_stat) {
    const char* dirs3[3] = {"MEAN_ONLINE", "VARIANCE_ONLINE", "SECOND_MOMENT_ONLINE"};
    for (const char* d : dirs3) {
      mkdirs(d);
      std::ofstream(std::string(d) + "/master_file.visit") << "!NBLOCKS 1" << std::endl;
    }
    // derived fields in the writer's attach order
    std::vector<std::string> dnames;
    for (const char* n : {"density", "mach", "vorticity", "entropy"})
      if (listed(n)) dnames.push_back(n);
    auto derived = [&](const std::string& n, int st, int k, int i) {
      const double* w = &snap[st][k][4 * (size_t)i];
      if (n == "density") return M.rho(w);
      if (n == "mach") return M.mach(w);
      if (n == "entropy") return M.entropy(w);
      return vsnap[st][k][i];
    };
    const size_t nd = dnames.size();
    std::vector<std::vector<double>> l1_mean, l1_sm;  // per stamp: T u v w p, then the derived fields
    for (int st = 0; st < nst; ++st) {
      std::vector<double> mean(4 * (size_t)nv, 0.0), m2(4 * (size_t)nv, 0.0), s2(4 * (size_t)nv, 0.0);
      std::vector<std::vector<double>> dm(nd, std::vector<double>(nv, 0.0)), dv = dm, ds = dm;
      int kk = 0;
      for (int k = 0; k < ns; ++k) {
        const auto& x = snap[st][k];
        if (x.empty()) continue;
        ++kk;
        for (size_t i = 0; i < x.size(); ++i) {
          if (kk == 1) { mean[i] = x[i]; s2[i] = x[i] * x[i]; m2[i] = 0.0; continue; }
          double Mk1 = mean[i];
          double Mk = x[i] * (1.0 / kk) + Mk1 * ((kk - 1.0) / kk);
          mean[i] = Mk;
          s2[i] += x[i] * x[i];
          m2[i] += (x[i] - Mk1) * (x[i] - Mk);
        }
        for (size_t q = 0; q < nd; ++q)
          for (int i = 0; i < nv; ++i) {
            const double val = derived(dnames[q], st, k, i);
            if (kk == 1) { dm[q][i] = val; ds[q][i] = val * val; dv[q][i] = 0.0; continue; }
            const double sMk1 = dm[q][i], sMk = sMk1 + (val - sMk1) / kk;
            dm[q][i] = sMk;
            ds[q][i] += val * val;
            dv[q][i] += (val - sMk1) * (val - sMk);
          }
      }
      if (kk < 2) continue;
      const double fact3 = 1.0 / ((double)kk - 1.0), fact4 = 1.0 / (double)kk;
      std::vector<double> var(m2.size()), sm(s2.size());
      for (size_t i = 0; i < m2.size(); ++i) { var[i] = m2[i] * fact3; sm[i] = s2[i] * fact4; }
      for (size_t q = 0; q < nd; ++q)
        for (int i = 0; i < nv; ++i) { dv[q][i] *= fact3; ds[q][i] *= fact4; }
      // L1 norms: sum over vertices of dual area x |value| (one partition)
      std::vector<double> a(5 + nd, 0.0), b(5 + nd, 0.0);
      for (int i = 0; i < nv; ++i) {
        const double area = dcarea[i];
        const int col[5] = {0, 1, 2, -1, 3};
        for (int c = 0; c < 5; ++c) {
          a[c] += col[c] < 0 ? 0.0 : area * std::abs(mean[4 * (size_t)i + col[c]]) / 1.0;
          b[c] += col[c] < 0 ? 0.0 : area * std::abs(sm[4 * (size_t)i + col[c]]) / 1.0;
        }
        for (size_t q = 0; q < nd; ++q) {
          a[5 + q] += area * std::abs(dm[q][i]) / 1.0;
          b[5 + q] += area * std::abs(ds[q][i]) / 1.0;
        }
      }
      l1_mean.push_back(a);
      l1_sm.push_back(b);
      std::vector<double> zero(8 * (size_t)nv, 0.0);
      tsm::Settings S0 = S;
      S0.write_variables.clear();
      const double ts = steady ? 0.0 : st * (S.p.final_time / S.n_time_stamps);
      const std::string names[3] = {stamp_name("mean", st), stamp_name("var", st), stamp_name("sec_mom", st)};
      const std::vector<double>* fields[3] = {&mean, &var, &sm};
      const std::vector<std::vector<double>>* dfields[3] = {&dm, &dv, &ds};
      for (int f = 0; f < 3; ++f) {
        const std::string file = std::string(dirs3[f]) + "/" + names[f] + (raw ? ".raw" : ".vtk");
        if (raw) {
          if (!write_raw(file, *fields[f], nv, ts)) {
            std::cerr << "ERROR --> cannot write " << file << std::endl;
            return 1;
          }
        } else {
          Attached more;
          for (size_t q = 0; q < nd; ++q) more.push_back({dnames[q], &(*dfields[f])[q]});
          write_vtk(file, P, *fields[f], zero, dcarea, M, S0, ts, more);
        }
        std::ofstream(std::string(dirs3[f]) + "/master_file.visit", std::ios::app)
            << names[f] << (raw ? ".raw" : ".vtk") << std::endl;
      }
    }
    // The reference names this directory but never sets its name (src/fv.h:47),
    // so its two files go to the file system root; here they go to
    // STAT_L1_ONLINE.  The "var" file holds the second moment's norm, as there.
    if (!l1_mean.empty()) {
      mkdirs("STAT_L1_ONLINE");
      const char* fn[2] = {"STAT_L1_ONLINE/mean_l1_norm.dat", "STAT_L1_ONLINE/var_l1_norm.dat"};
      const std::vector<std::vector<double>>* rows[2] = {&l1_mean, &l1_sm};
      for (int f = 0; f < 2; ++f) {
        std::ofstream o(fn[f]);
        o << " TIME TEMPERATURE U V W PRESSURE";
        for (const char* n : {"density", "mach", "vorticity", "entropy"})  // the reference's header order
          if (listed(n)) o << " " << (std::string(n) == "density" ? "DENSITY" : std::string(n) == "mach" ? "MACH" : std::string(n) == "vorticity" ? "VORTICITY" : "ENTROPY");
        o << "\n";
        for (size_t t = 0; t < rows[f]->size(); ++t) {
          const auto& r = (*rows[f])[t];
          o << t;
          for (int c = 0; c < 5; ++c) o << " " << r[c];
          for (const char* n : {"density", "mach", "vorticity", "entropy"})
            for (size_t q = 0; q < nd; ++q)
              if (dnames[q] == n) o << " " << r[5 + q];
          o << "\n";
        }
      }
    }
  }
  if (!writer.finish()) {
    std::cerr << "ERROR --> an output file could not be written" << std::endl;
    tsm_destroy(ctx);
    return 1;
  }
  tsm_destroy(ctx);
  return 0;
}
