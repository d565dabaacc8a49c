// cac_host.cpp -- the host-side C++ layer over the C ABI: data-file reader,
// device context, Pair / Neighbor / FixNVE / Verlet mirrors.
#include "cac_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace cac_host {

// ------------------------------------------------------------------ read_data (read_data.cpp:338-375)

namespace {

std::string strip(const std::string& s) {
  const size_t h = s.find('#');
  std::string t = h == std::string::npos ? s : s.substr(0, h);
  const size_t a = t.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = t.find_last_not_of(" \t\r\n");
  return t.substr(a, b - a + 1);
}

bool is_section(const std::string& s) {
  static const char* names[] = {"Masses", "Element Types", "Atoms", "Velocities", "Elements", "Nodes",
                                "Node Velocities"};
  for (const char* n : names)
    if (s == n || s.rfind(std::string(n) + " ", 0) == 0) return true;
  return false;
}

[[noreturn]] void bad(const std::string& path, const std::string& what) {
  throw Error(CAC_HIP_EARG, "read_data " + path + ": " + what);
}

}  // namespace

void read_data(const std::string& path, System& sys) {
  std::ifstream in(path);
  if (!in) bad(path, "cannot open");
  std::vector<std::string> lines;
  for (std::string l; std::getline(in, l);) lines.push_back(l);
  size_t i = 1;  // the first line is a title
  int na = 0, ne = 0, ntypes = 0;
  for (; i < lines.size(); i++) {
    const std::string s = strip(lines[i]);
    if (s.empty()) continue;
    if (is_section(s)) break;
    std::istringstream w(s);
    std::vector<std::string> t;
    for (std::string x; w >> x;) t.push_back(x);
    if (t.size() >= 2 && t[1] == "atoms") na = std::atoi(t[0].c_str());
    else if (t.size() >= 2 && t[1] == "elements") ne = std::atoi(t[0].c_str());
    else if (t.size() >= 3 && t[1] == "atom" && t[2] == "types") ntypes = std::atoi(t[0].c_str());
    else if (t.size() == 4 && t[2].size() == 3 && t[2].substr(1) == "lo") {
      const int d = t[2][0] - 'x';
      if (d < 0 || d > 2) bad(path, "bad box line: " + s);
      sys.boxlo[d] = std::atof(t[0].c_str());
      sys.boxhi[d] = std::atof(t[1].c_str());
    }
  }
  sys.mass.assign(ntypes, 0.0);
  Atoms& at = sys.atom;
  Elements& el = sys.element;
  at = Atoms();
  el = Elements();
  at.n = na;
  at.x.assign(3 * (size_t)na, 0.0);
  at.v.assign(3 * (size_t)na, 0.0);
  at.f.assign(3 * (size_t)na, 0.0);
  at.type.assign(na, 1);
  at.tag.assign(na, 0);
  el.n = ne;
  el.nodex.assign(24 * (size_t)ne, 0.0);
  el.nodev.assign(24 * (size_t)ne, 0.0);
  el.nodef.assign(24 * (size_t)ne, 0.0);
  el.gaussf.assign(24 * (size_t)ne, 0.0);
  el.etype.assign(ne, 1);
  el.ctype.assign(ne, 1);
  el.tag.assign(ne, 0);
  std::unordered_map<long, int> aidx, eidx;
  int nseen_a = 0, nseen_e = 0;
  std::string sec;
  for (; i < lines.size(); i++) {
    const std::string s = strip(lines[i]);
    if (s.empty()) continue;
    if (is_section(s)) {
      sec = s;
      continue;
    }
    std::istringstream w(s);
    if (sec == "Masses") {
      int t;
      double m;
      if (!(w >> t >> m) || t < 1 || t > ntypes) bad(path, "bad Masses line: " + s);
      sys.mass[t - 1] = m;
    } else if (sec == "Element Types") {
      int it, apc, nc[3], ng[3];
      std::string shape;
      if (!(w >> it >> shape >> apc >> nc[0] >> nc[1] >> nc[2] >> ng[0] >> ng[1] >> ng[2]))
        bad(path, "bad Element Types line: " + s);
      if (shape != "Hex" || apc != 1 || ng[0] != 2 || ng[1] != 2 || ng[2] != 2)
        bad(path, "this build takes Hex elements, 1 atom per unit cell, 2 2 2 Gauss cells");
      if (el.ncell[0] && (el.ncell[0] != nc[0] || el.ncell[1] != nc[1] || el.ncell[2] != nc[2]))
        bad(path, "one element type per model");
      for (int d = 0; d < 3; d++) el.ncell[d] = nc[d];
    } else if (sec == "Atoms") {
      long id;
      int t;
      double x, y, z;
      if (!(w >> id >> t >> x >> y >> z)) bad(path, "bad Atoms line: " + s);
      if (nseen_a >= na || aidx.count(id)) bad(path, "atom ids: more atoms than declared or a repeat");
      const int k = nseen_a++;
      aidx[id] = k;
      at.tag[k] = (int)id;
      at.type[k] = t;
      at.x[3 * (size_t)k] = x;
      at.x[3 * (size_t)k + 1] = y;
      at.x[3 * (size_t)k + 2] = z;
    } else if (sec == "Velocities") {
      long id;
      double vx, vy, vz;
      if (!(w >> id >> vx >> vy >> vz)) bad(path, "bad Velocities line: " + s);
      auto it = aidx.find(id);
      if (it == aidx.end()) bad(path, "velocity of an unknown atom");
      at.v[3 * (size_t)it->second] = vx;
      at.v[3 * (size_t)it->second + 1] = vy;
      at.v[3 * (size_t)it->second + 2] = vz;
    } else if (sec == "Elements") {
      long id;
      int et;
      double xc, yc, zc;
      if (!(w >> id >> et >> xc >> yc >> zc)) bad(path, "bad Elements line: " + s);
      if (nseen_e >= ne || eidx.count(id)) bad(path, "element ids: more elements than declared or a repeat");
      const int k = nseen_e++;
      eidx[id] = k;
      el.tag[k] = (int)id;
      el.etype[k] = et;
    } else if (sec == "Nodes" || sec == "Node Velocities") {
      long id;
      int inode, ibasis, ctype = 1;
      double a, b, c;
      const bool pos = sec == "Nodes";
      if (!(w >> id >> inode >> ibasis)) bad(path, "bad " + sec + " line: " + s);
      if (pos && !(w >> ctype)) bad(path, "bad Nodes line: " + s);
      if (!(w >> a >> b >> c)) bad(path, "bad " + sec + " line: " + s);
      auto it = eidx.find(id);
      if (it == eidx.end()) bad(path, "node of an unknown element");
      if (inode < 1 || inode > 8 || ibasis != 1) bad(path, "node index out of range");
      double* dst = (pos ? el.nodex.data() : el.nodev.data()) + ((size_t)it->second * 8 + inode - 1) * 3;
      dst[0] = a;
      dst[1] = b;
      dst[2] = c;
      if (pos) el.ctype[it->second] = ctype;
    }
  }
  if (nseen_a != na) bad(path, "fewer atoms than declared");
  if (nseen_e != ne) bad(path, "fewer elements than declared");
  if (ne && !el.ncell[0]) bad(path, "elements but no Element Types section");
}

// ------------------------------------------------------------------ write_data (write_data.cpp:188-600)

void write_data_file(const std::string& path, const System& sys, long ntimestep, bool velocities) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) throw Error(CAC_HIP_EARG, "Cannot open data file " + path);
  const Atoms& at = sys.atom;
  const Elements& el = sys.element;
  std::fprintf(fp, "CAC data file via write_data, timestep = %ld\n\n", ntimestep);
  if (at.n) std::fprintf(fp, "%d atoms\n", at.n);
  if (el.n) std::fprintf(fp, "%d elements\n", el.n);
  std::fprintf(fp, "%d atom types\n", (int)sys.mass.size());
  if (el.n) std::fprintf(fp, "%d element types\n", 1);
  std::fprintf(fp, "\n");
  const char* ax[3] = {"x", "y", "z"};
  for (int d = 0; d < 3; d++)
    std::fprintf(fp, "%-1.16e %-1.16e %slo %shi\n", sys.boxlo[d], sys.boxhi[d], ax[d], ax[d]);
  if (!sys.mass.empty()) {
    std::fprintf(fp, "\nMasses\n\n");
    for (size_t t = 0; t < sys.mass.size(); t++) std::fprintf(fp, "%d %g\n", (int)t + 1, sys.mass[t]);
  }
  if (el.n) {
    // one Hex type, one atom per unit cell, 2x2x2 Gauss cells (fewer along
    // an axis of one unit cell), one point per cell
    std::fprintf(fp, "\nElement Types\n\n");
    std::fprintf(fp, "%d %s %d %d %d %d %d %d %d %d %d %d\n", 1, "Hex", 1, el.ncell[0], el.ncell[1], el.ncell[2],
                 std::min(2, el.ncell[0]), std::min(2, el.ncell[1]), std::min(2, el.ncell[2]), 1, 1, 1);
  }
  if (at.n) {
    std::fprintf(fp, "\nAtoms\n\n");
    for (int i = 0; i < at.n; i++) {
      const double* x = &at.x[3 * (size_t)i];
      std::fprintf(fp, "%d %d %-1.16e %-1.16e %-1.16e\n", at.tag[i], at.type[i], x[0], x[1], x[2]);
    }
    if (velocities) {
      std::fprintf(fp, "\nVelocities\n\n");
      for (int i = 0; i < at.n; i++) {
        const double* v = &at.v[3 * (size_t)i];
        std::fprintf(fp, "%d %-1.16e %-1.16e %-1.16e\n", at.tag[i], v[0], v[1], v[2]);
      }
    }
  }
  if (el.n) {
    std::fprintf(fp, "\nElements\n\n");
    for (int e = 0; e < el.n; e++) {
      double c[3] = {0, 0, 0};
      for (int k = 0; k < 8; k++)
        for (int d = 0; d < 3; d++) c[d] += el.nodex[((size_t)e * 8 + k) * 3 + d];
      std::fprintf(fp, "%d %d %-1.16e %-1.16e %-1.16e\n", el.tag[e], el.etype[e], c[0] / 8, c[1] / 8, c[2] / 8);
    }
    std::fprintf(fp, "\nNodes\n\n");
    for (int e = 0; e < el.n; e++)
      for (int k = 0; k < 8; k++) {
        const double* x = &el.nodex[((size_t)e * 8 + k) * 3];
        std::fprintf(fp, "%d %d %d %d %-1.16e %-1.16e %-1.16e\n", el.tag[e], k + 1, 1, el.ctype[e], x[0], x[1], x[2]);
      }
    if (velocities) {
      std::fprintf(fp, "\nNode Velocities\n\n");
      for (int e = 0; e < el.n; e++)
        for (int k = 0; k < 8; k++) {
          const double* v = &el.nodev[((size_t)e * 8 + k) * 3];
          std::fprintf(fp, "%d %d %d %-1.16e %-1.16e %-1.16e\n", el.tag[e], k + 1, 1, v[0], v[1], v[2]);
        }
    }
  }
  const bool ok = std::ferror(fp) == 0;
  if (std::fclose(fp) != 0 || !ok) throw Error(CAC_HIP_EARG, "write_data " + path + ": write failed");
}

void write_data(const std::string& path, Device& dev, long ntimestep, bool velocities) {
  dev.pbc();
  dev.download_state();
  write_data_file(path, dev.system(), ntimestep, velocities);
}

// ------------------------------------------------------------------ device context

Device::Device(System& sys, int device) : sys_(sys) {
  const int rc = cac_hip_create(&ctx_, device, sys.units);
  if (rc) throw Error(rc, "cac_hip_create failed (no HIP device?)");
  upload();
}

Device::~Device() {
  if (ctx_) cac_hip_destroy(ctx_);
}

void Device::check(int rc) const {
  if (rc) throw Error(rc, cac_hip_last_error(ctx_));
}

void Device::upload() {
  System& s = sys_;
  check(cac_hip_set_box(ctx_, s.boxlo, s.boxhi));
  check(cac_hip_set_masses(ctx_, (int)s.mass.size(), s.mass.data()));
  if (s.element.n)
    check(cac_hip_add_etype(ctx_, 1, s.element.ncell[0], s.element.ncell[1], s.element.ncell[2], 2, 2, 2));
  check(cac_hip_upload_atoms(ctx_, s.atom.n, s.atom.x.data(), s.atom.v.data(), s.atom.type.data()));
  if (s.element.n)
    check(cac_hip_upload_elements(ctx_, s.element.n, s.element.etype.data(), s.element.ctype.data(),
                                  s.element.nodex.data(), s.element.nodev.data()));
}

void Device::sync_positions() {
  check(cac_hip_sync_positions(ctx_, sys_.atom.n ? sys_.atom.x.data() : nullptr,
                               sys_.element.n ? sys_.element.nodex.data() : nullptr));
}

void Device::download_state() {
  if (sys_.atom.n) check(cac_hip_download_atoms(ctx_, sys_.atom.x.data(), sys_.atom.v.data(), nullptr));
  if (sys_.element.n)
    check(cac_hip_download_elements(ctx_, sys_.element.nodex.data(), sys_.element.nodev.data(), nullptr, nullptr));
}

void Device::pbc() { check(cac_hip_pbc(ctx_)); }

void Device::download_forces() {
  check(cac_hip_download_forces(ctx_, sys_.atom.n ? sys_.atom.f.data() : nullptr,
                                sys_.element.n ? sys_.element.nodef.data() : nullptr,
                                sys_.element.n ? sys_.element.gaussf.data() : nullptr));
}

// ------------------------------------------------------------------ pair styles

void Pair::compute(int eflag, int vflag) {
  dev_.check(cac_hip_compute(dev_.ctx(), eflag, vflag));
  dev_.check(cac_hip_get_energy(dev_.ctx(), &eng_vdwl, virial));
  const System& s = dev_.system();
  const bool ea = eflag & 2, va = vflag & 12;  // ENERGY_ATOM; VIRIAL_ATOM | VIRIAL_CENTROID
  if (!ea && !va) return;
  if (ea) {
    eatom.assign((size_t)s.atom.n, 0.0);
    enode.assign((size_t)s.element.n * 8, 0.0);
  }
  if (va) {
    vatom.assign((size_t)s.atom.n * 6, 0.0);
    vnode.assign((size_t)s.element.n * 48, 0.0);
  }
  dev_.check(cac_hip_get_peratom(dev_.ctx(), ea ? eatom.data() : nullptr, ea ? enode.data() : nullptr,
                                 va ? vatom.data() : nullptr, va ? vnode.data() : nullptr));
}

void PairEAMAlloyHIP::settings(int narg, char**) {
  if (narg != 0) throw Error(CAC_HIP_EARG, "Illegal pair_style command");
}

// pair_coeff * * file.setfl Element (pair_eam_alloy.cpp:56-95): one element type
void PairEAMAlloyHIP::coeff(int narg, char** arg) {
  if (narg < 4 || std::strcmp(arg[0], "*") != 0 || std::strcmp(arg[1], "*") != 0)
    throw Error(CAC_HIP_EARG, "Incorrect args for pair coefficients");
  dev_.check(cac_hip_pair_eam_setfl(dev_.ctx(), arg[2]));
  // the cutoff is the fifth field of the fifth line (pair_eam_alloy.cpp:120-135)
  std::ifstream in(arg[2]);
  std::string l;
  for (int k = 0; k < 5 && std::getline(in, l); k++) {
  }
  std::istringstream w(l);
  int nrho, nr;
  double drho, dr;
  w >> nrho >> drho >> nr >> dr >> cutmax;
}

double PairEAMAlloyHIP::init_one(int, int) { return cutmax; }

void PairLJCutHIP::settings(int narg, char** arg) {
  if (narg != 1) throw Error(CAC_HIP_EARG, "Illegal pair_style command");
  cut_global = std::atof(arg[0]);
  cut = cut_global;
}

// pair_coeff 1 1 epsilon sigma [cut] (pair_lj_cut.cpp:105-136), one atom type
void PairLJCutHIP::coeff(int narg, char** arg) {
  if (narg < 4 || narg > 5) throw Error(CAC_HIP_EARG, "Incorrect args for pair coefficients");
  epsilon = std::atof(arg[2]);
  sigma = std::atof(arg[3]);
  cut = narg == 5 ? std::atof(arg[4]) : cut_global;
  dev_.check(cac_hip_pair_lj(dev_.ctx(), epsilon, sigma, cut));
}

double PairLJCutHIP::init_one(int, int) { return cut; }

// ------------------------------------------------------------------ neighbor, fix, verlet

void NeighborHIP::init() { dev_.check(cac_hip_neighbor_settings(dev_.ctx(), skin, every, delay, check, oneatom)); }

int NeighborHIP::decide() {
  int flag = 0;
  dev_.check(cac_hip_neigh_decide(dev_.ctx(), &flag));
  return flag;
}

void NeighborHIP::build() { dev_.check(cac_hip_neigh_build(dev_.ctx())); }

int NeighborHIP::ncalls() const {
  int a = 0, b = 0;
  dev_.check(cac_hip_neigh_stats(dev_.ctx(), &a, &b));
  return a;
}

int NeighborHIP::ndanger() const {
  int a = 0, b = 0;
  dev_.check(cac_hip_neigh_stats(dev_.ctx(), &a, &b));
  return b;
}

void FixNVEHIP::initial_integrate(int) { dev_.check(cac_hip_nve_initial(dev_.ctx(), dt_)); }

void FixNVEHIP::final_integrate() { dev_.check(cac_hip_nve_final(dev_.ctx(), dt_)); }

ThermoRow VerletHIP::thermo(long step) {
  ThermoRow r{};
  r.step = step;
  dev_.check(cac_hip_thermo(dev_.ctx(), &r.pe, &r.ke, &r.temp, &r.press));
  r.etotal = r.pe + r.ke;
  return r;
}

void VerletHIP::setup() {
  neigh_.init();
  pair_.init_style();
  dev_.check(cac_hip_setup(dev_.ctx(), 1, 1));
  dev_.check(cac_hip_get_energy(dev_.ctx(), &pair_.eng_vdwl, pair_.virial));
}

std::vector<ThermoRow> VerletHIP::run(int nsteps, int thermo_every) {
  std::vector<ThermoRow> rows;
  rows.push_back(thermo(ntimestep));
  for (int s = 0; s < nsteps; s++) {
    ntimestep++;
    const bool ev = (thermo_every > 0 && ntimestep % thermo_every == 0) || s == nsteps - 1;
    fix_.initial_integrate(0);
    if (neigh_.decide()) neigh_.build();
    pair_.compute(ev ? 1 : 0, ev ? 1 : 0);
    fix_.final_integrate();
    if (ev) rows.push_back(thermo(ntimestep));
  }
  return rows;
}

std::vector<ThermoRow> VerletHIP::run_resident(int nsteps, double dt, int thermo_every) {
  const int cap = (thermo_every > 0 ? nsteps / thermo_every : 0) + 3;
  std::vector<double> th((size_t)5 * cap, 0.0);
  dev_.check(cac_hip_run(dev_.ctx(), nsteps, dt, thermo_every, th.data()));
  int used = 1 + (thermo_every > 0 ? nsteps / thermo_every : 0);
  if (thermo_every <= 0 || nsteps % thermo_every) used++;
  std::vector<ThermoRow> rows;
  const System& s = dev_.system();
  int nucell = s.element.ncell[0] * s.element.ncell[1] * s.element.ncell[2];
  const double dof = 3.0 * ((double)s.atom.n + (double)nucell * s.element.n);
  const double boltz = s.units == CAC_HIP_UNITS_METAL ? 8.617343e-5 : 1.0;
  for (int r = 0; r < used; r++) {
    const double* t = &th[(size_t)5 * r];
    ThermoRow row{(long)t[0], t[1], t[2], t[3], t[4], dof > 0 ? t[2] / (0.5 * dof * boltz) : 0.0};
    rows.push_back(row);
  }
  ntimestep += nsteps;
  return rows;
}

}  // namespace cac_host
