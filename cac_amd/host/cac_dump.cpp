// cac_dump.cpp -- output taps of the host layer: "dump custom" snapshots and
// write_dump_neighbor_debug, in the reference's text formats, so that device
// forces and lists can be diffed against the reference's own output files.
//
//   DumpCustom                dump_custom.cpp:30-121 (arguments), 447-640
//                             (columns), 655-920 (values), 1260-1420
//                             (atoms then virtual atoms), 1672-1719 (lines and
//                             header); file naming as dump.cpp:98-136, 233-280
//   write_dump_neighbor_debug write_dump_neighbor_debug.cpp:41-409
//
// Values come from the host copy of the model (System), which the caller
// refreshes from the device (Device::download_state / download_forces); the
// neighbor rows come from cac_hip_neigh_export.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cac_host.hpp"

namespace cac_host {

namespace {

[[noreturn]] void fail(const std::string& m) { throw Error(CAC_HIP_EARG, m); }

// Hex node natural coordinates in the reference's node order
// (element_vec_cac.cpp:1756-1764): (-,-,-) (+,-,-) (+,+,-) (-,+,-) (-,-,+) ...
const int HEX_SIGN[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                            {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// Trilinear shape functions at unit cell u of an nx x ny x nz element, unit
// cells numbered x-major (u = i*ny*nz + j*nz + k), natural coordinate of
// index i = (2 / (n - 1)) * i - 1 (element_vec_cac.cpp:1881-1900,
// element.cpp:2993-3001).
void hex_shape(const int* nc, int u, double* N) {
  const int i = u / (nc[1] * nc[2]), j = (u / nc[2]) % nc[1], k = u % nc[2];
  const double p[3] = {2.0 / (nc[0] - 1) * i - 1.0, 2.0 / (nc[1] - 1) * j - 1.0, 2.0 / (nc[2] - 1) * k - 1.0};
  for (int n = 0; n < 8; n++)
    N[n] = 0.125 * (1 + HEX_SIGN[n][0] * p[0]) * (1 + HEX_SIGN[n][1] * p[1]) * (1 + HEX_SIGN[n][2] * p[2]);
}

// the unit cell of node n, and of Gauss cell g (2x2x2 Gauss cells sit on the
// element corners; Gauss cells are numbered x-major over the corners)
int node_ucell(const int* nc, int n) {
  const int i = HEX_SIGN[n][0] > 0 ? nc[0] - 1 : 0, j = HEX_SIGN[n][1] > 0 ? nc[1] - 1 : 0,
            k = HEX_SIGN[n][2] > 0 ? nc[2] - 1 : 0;
  return (i * nc[1] + j) * nc[2] + k;
}
int gauss_ucell(const int* nc, int g) {
  const int i = (g >> 2) & 1 ? nc[0] - 1 : 0, j = (g >> 1) & 1 ? nc[1] - 1 : 0, k = g & 1 ? nc[2] - 1 : 0;
  return (i * nc[1] + j) * nc[2] + k;
}
int ucell_node(const int* nc, int u) {
  for (int n = 0; n < 8; n++)
    if (node_ucell(nc, n) == u) return n;
  return -1;
}
int ucell_gauss(const int* nc, int u) {
  for (int g = 0; g < 8; g++)
    if (gauss_ucell(nc, g) == u) return g;
  return -1;
}

// sum_n N[n] * a[e][n][d], in node order (ElementVec::interpolate)
double interp(const std::vector<double>& a, int e, const double* N, int d) {
  double s = 0.0;
  for (int n = 0; n < 8; n++) s += N[n] * a[((size_t)e * 8 + n) * 3 + d];
  return s;
}

// Domain::remap for an orthogonal periodic box (domain.cpp:1543-1580)
void remap(const System& s, double* x) {
  for (int d = 0; d < 3; d++) {
    const double lo = s.boxlo[d], hi = s.boxhi[d], prd = hi - lo;
    while (x[d] < lo) x[d] += prd;
    while (x[d] >= hi) x[d] -= prd;
    x[d] = std::max(x[d], lo);
  }
}

std::string step_name(const std::string& pattern, long step, const std::string& suffix) {
  const size_t star = pattern.find('*');
  if (star == std::string::npos) return pattern + suffix;
  return pattern.substr(0, star) + std::to_string(step) + pattern.substr(star + 1) + suffix;
}

void box_header(FILE* fp, const System& s) {
  std::fprintf(fp, "ITEM: BOX BOUNDS pp pp pp\n");
  for (int d = 0; d < 3; d++) std::fprintf(fp, "%-1.16e %-1.16e\n", s.boxlo[d], s.boxhi[d]);
}

int max_atom_tag(const System& s) {
  int m = 0;
  for (int t : s.atom.tag) m = std::max(m, t);
  return m;
}

}  // namespace

// ------------------------------------------------------------------ dump custom

DumpCustom::DumpCustom(int narg, char** arg) {
  // dump ID group custom N file [format F] [average yes|no] [wrap yes|no] [precision high|normal] col ...
  if (narg < 6) fail("Illegal dump custom/tecplot command");
  id = arg[0];
  if (std::strcmp(arg[1], "all") != 0) fail("dump custom: only group all is available");
  if (std::strcmp(arg[2], "custom") != 0) fail("dump: only style custom is available");
  nevery = std::atoi(arg[3]);
  if (nevery <= 0) fail("Illegal dump custom command");
  file = arg[4];
  int iarg = 5, cols_start = -1;
  for (; iarg < narg; iarg += 2) {
    const std::string a = arg[iarg];
    const bool has_value = iarg + 1 < narg;
    if (a == "format") {
      if (!has_value) fail("Illegal dump custom/atom command: missing style value");
      const std::string f = arg[iarg + 1];
      if (f == "lammps/full/map") style = FULLMAP;
      else if (f == "lammps/surface") style = SURFACE;
      else if (f == "lammps/gauss/point") style = GAUSSPOINT;
      else if (f == "lammps/node") style = NODE;
      else if (f == "lammps/center") style = CENTER;
      else if (f.rfind("tecplot/", 0) == 0) fail("dump custom: tecplot formats are not available in this build");
      else fail("Illegal dump custom/atom command: unknown style");
    } else if (a == "average") {
      if (!has_value) fail("Illegal dump custom/tecplot command");
      const std::string v = arg[iarg + 1];
      if (v != "yes" && v != "no") fail("Illegal dump custom/tecplot command: average yes/no expected");
    } else if (a == "wrap" || a == "precision") {
      if (!has_value) fail("Illegal dump atom command");
      const std::string v = arg[iarg + 1];
      if (a == "wrap") {
        if (v != "yes" && v != "no") fail("Illegal dump custom command");
        wrap = v == "yes";
      } else {
        if (v != "high" && v != "normal") fail("Illegal dump custom command");
        precision_high = v == "high";
      }
      // the reference advances past one more keyword pair after wrap and
      // precision (dump_custom.cpp:101-112): the pair that follows is skipped
      iarg += 2;
    } else {
      cols_start = iarg;
      break;
    }
  }
  if (cols_start < 0 || cols_start >= narg) fail("No column keywords found for dump custom/atom command");
  static const struct {
    const char* key;
    Col col;
    bool is_int;
  } table[] = {{"id", ID, true},       {"type", TYPE, true},     {"x", X, false},         {"y", Y, false},
               {"z", Z, false},        {"vx", VX, false},        {"vy", VY, false},       {"vz", VZ, false},
               {"fx", FX, false},      {"fy", FY, false},        {"fz", FZ, false},       {"mass", MASS, false},
               {"proc", PROC, true},   {"procp1", PROCP1, true}, {"eid", EID, true},      {"etype", ETYPE, true},
               {"ibasis", IBASIS, true}, {"igauss", IGAUSS, true}, {"inode", INODE, true}, {"iucell", IUCELL, true}};
  for (int k = cols_start; k < narg; k++) {
    const std::string key = arg[k];
    bool found = false;
    for (const auto& t : table)
      if (key == t.key) {
        cols.push_back(t.col);
        col_int.push_back(t.is_int);
        found = true;
        break;
      }
    if (!found) {
      if (key == "ix" || key == "iy" || key == "iz" || key == "xu" || key == "yu" || key == "zu")
        fail("dump custom: image-flag columns need image counts, which this build does not track");
      if (key.rfind("c_", 0) == 0 || key.rfind("f_", 0) == 0)
        fail("dump custom: compute and fix columns are not available in this build");
      fail("Dump custom: unknown column keyword");
    }
    if (!columns.empty()) columns += " ";
    columns += key;
  }
  // precision high (dump_custom.cpp:648-652) is meant to print doubles as
  // %-1.16e; this tap applies it
}

void DumpCustom::row(std::vector<double>& buf, const System& s, int kind, int i, int ibasis, int iucell, int igcell,
                     int inode, long vid, const double* x) const {
  // kind 0: atom i; kind 1: virtual atom of element i; kind 2: element centre
  const Elements& el = s.element;
  double N[8];
  if (kind == 1) hex_shape(el.ncell, iucell, N);
  for (Col c : cols) {
    double v = 0.0;
    if (kind == 0) {
      switch (c) {
        case ID: v = s.atom.tag[i]; break;
        case TYPE: v = s.atom.type[i]; break;
        case X: case Y: case Z: v = x[c - X]; break;
        case VX: case VY: case VZ: v = s.atom.v.empty() ? 0.0 : s.atom.v[(size_t)i * 3 + (c - VX)]; break;
        case FX: case FY: case FZ: v = s.atom.f.empty() ? 0.0 : s.atom.f[(size_t)i * 3 + (c - FX)]; break;
        case MASS: v = s.mass[s.atom.type[i] - 1]; break;
        case PROC: v = 0; break;
        case PROCP1: v = 1; break;
        default: v = -1.0;  // debug columns of real atoms
      }
    } else if (kind == 1) {
      switch (c) {
        case ID: v = (double)vid; break;
        case TYPE: v = el.ctype[i]; break;
        case X: case Y: case Z: v = x[c - X]; break;
        case VX: case VY: case VZ: v = el.nodev.empty() ? 0.0 : interp(el.nodev, i, N, c - VX); break;
        case FX: case FY: case FZ: v = el.nodef.empty() ? 0.0 : interp(el.nodef, i, N, c - FX); break;
        case MASS: v = s.mass[el.ctype[i] - 1]; break;
        case PROC: v = 0; break;
        case PROCP1: v = 1; break;
        case EID: v = el.tag[i]; break;
        case ETYPE: v = el.etype[i]; break;
        case IBASIS: v = ibasis; break;
        case IGAUSS: v = igcell; break;
        case INODE: v = inode; break;
        case IUCELL: v = iucell; break;
      }
    } else {
      switch (c) {
        case ID: v = (double)vid; break;
        case TYPE: v = 0; break;
        case X: case Y: case Z: v = x[c - X]; break;
        case PROC: v = 0; break;
        case PROCP1: v = 1; break;
        case EID: v = el.tag[i]; break;
        case ETYPE: v = el.etype[i]; break;
        case IBASIS: case IGAUSS: case INODE: case IUCELL: v = -1; break;
        default: v = 0.0;
      }
    }
    buf.push_back(v);
  }
}

long DumpCustom::count(const System& s) const {
  const int* nc = s.element.ncell;
  const long nucell = (long)nc[0] * nc[1] * nc[2];
  long per = 0;
  switch (style) {
    case FULLMAP: per = nucell; break;
    case SURFACE: per = nucell - (long)std::max(nc[0] - 2, 0) * std::max(nc[1] - 2, 0) * std::max(nc[2] - 2, 0); break;
    case GAUSSPOINT: case NODE: per = 8; break;
    case CENTER: per = 1; break;
  }
  return s.atom.n + per * s.element.n;
}

void DumpCustom::write(const System& s, long ntimestep) {
  const bool multifile = file.find('*') != std::string::npos;
  FILE* fp = nullptr;
  if (multifile) fp = std::fopen(step_name(file, ntimestep, "").c_str(), "w");
  else {
    fp = std::fopen(file.c_str(), opened_ ? "a" : "w");
    opened_ = true;
  }
  if (!fp) fail("Cannot open dump file " + file);
  std::fprintf(fp, "ITEM: TIMESTEP\n%ld\nITEM: NUMBER OF ATOMS\n%ld\n", ntimestep, count(s));
  box_header(fp, s);
  std::fprintf(fp, "ITEM: ATOMS %s\n", columns.c_str());
  const Elements& el = s.element;
  const int* nc = el.ncell;
  const int nucell = nc[0] * nc[1] * nc[2];
  const long itag = (long)max_atom_tag(s) + 1;
  std::vector<double> buf;
  buf.reserve(cols.size());
  auto emit = [&]() {
    for (size_t k = 0; k < cols.size(); k++) {
      if (k) std::fputc(' ', fp);
      if (col_int[k]) std::fprintf(fp, "%d", (int)buf[k]);
      else std::fprintf(fp, precision_high ? "%-1.16e" : "%g", buf[k]);
    }
    std::fputc('\n', fp);
    buf.clear();
  };
  for (int i = 0; i < s.atom.n; i++) {
    row(buf, s, 0, i, -1, -1, -1, -1, 0, &s.atom.x[(size_t)i * 3]);
    emit();
  }
  double x[3], N[8];
  for (int i = 0; i < el.n; i++) {
    const long et = el.tag[i];
    if (style == CENTER) {
      for (int d = 0; d < 3; d++) {
        double c = 0.0;
        for (int n = 0; n < 8; n++) c += el.nodex[((size_t)i * 8 + n) * 3 + d];
        x[d] = c / 8;
      }
      row(buf, s, 2, i, -1, -1, -1, -1, itag + et - 1, x);
      emit();
    } else if (style == GAUSSPOINT) {
      for (int g = 0; g < 8; g++) {
        const int u = gauss_ucell(nc, g);
        hex_shape(nc, u, N);
        for (int d = 0; d < 3; d++) x[d] = interp(el.nodex, i, N, d);
        if (wrap) remap(s, x);
        row(buf, s, 1, i, 0, u, g, ucell_node(nc, u), itag + (et - 1) * 8 + g, x);
        emit();
      }
    } else if (style == NODE) {
      for (int n = 0; n < 8; n++) {
        const int u = node_ucell(nc, n);
        for (int d = 0; d < 3; d++) x[d] = el.nodex[((size_t)i * 8 + n) * 3 + d];
        if (wrap) remap(s, x);
        row(buf, s, 1, i, 0, u, ucell_gauss(nc, u), n, itag + (et - 1) * 8 + n, x);
        emit();
      }
    } else {
      for (int u = 0; u < nucell; u++) {
        if (style == SURFACE) {
          const int a = u / (nc[1] * nc[2]), b = (u / nc[2]) % nc[1], c = u % nc[2];
          if (a != 0 && a != nc[0] - 1 && b != 0 && b != nc[1] - 1 && c != 0 && c != nc[2] - 1) continue;
        }
        hex_shape(nc, u, N);
        for (int d = 0; d < 3; d++) x[d] = interp(el.nodex, i, N, d);
        if (wrap) remap(s, x);
        row(buf, s, 1, i, 0, u, ucell_gauss(nc, u), ucell_node(nc, u), itag + (et - 1) * (long)nucell + u, x);
        emit();
      }
    }
  }
  std::fclose(fp);
}

// ------------------------------------------------------------------ neighbor rows

NeighborRows export_neighbor_rows(Device& dev) {
  long n = 0;
  dev.check(cac_hip_neigh_export(dev.ctx(), &n, nullptr, nullptr, nullptr, nullptr, nullptr));
  std::vector<int> row(n), kind(n), owner(n), ucell(n), img(3 * (size_t)n);
  if (n) dev.check(cac_hip_neigh_export(dev.ctx(), &n, row.data(), kind.data(), owner.data(), ucell.data(), img.data()));
  const System& s = dev.system();
  const int P = s.atom.n + 8 * s.element.n;
  NeighborRows nr;
  nr.first.assign((size_t)P + 1, 0);
  for (long k = 0; k < n; k++) nr.first[(size_t)row[k] + 1]++;
  for (int p = 0; p < P; p++) nr.first[(size_t)p + 1] += nr.first[(size_t)p];
  nr.kind.resize(n);
  nr.owner.resize(n);
  nr.ucell.resize(n);
  nr.img.resize(3 * (size_t)n);
  std::vector<long> fill(nr.first.begin(), nr.first.end() - 1);
  for (long k = 0; k < n; k++) {  // export order within a row is kept
    const long q = fill[(size_t)row[k]]++;
    nr.kind[q] = kind[k];
    nr.owner[q] = owner[k];
    nr.ucell[q] = ucell[k];
    for (int d = 0; d < 3; d++) nr.img[3 * (size_t)q + d] = img[3 * (size_t)k + d];
  }
  return nr;
}

// ------------------------------------------------------------------ write_dump_neighbor_debug

std::string write_dump_neighbor_debug(const System& s, const NeighborRows& nl, int narg, char** arg, long ntimestep) {
  // write_dump_neighbor_debug file [gauss id ibasis igcell | atom id]
  if (narg < 1) fail("Illegal write_dump_neighbor_debug command");
  int single = 0;
  long id = 0;
  int ibasis = 0, igcell = 0;
  if (narg > 1) {
    const std::string m = arg[1];
    if (m == "gauss") {
      if (narg != 5) fail("Illegal write_dump_neighbor_debug command");
      single = 1;
      id = std::atol(arg[2]);
      ibasis = std::atoi(arg[3]);
      igcell = std::atoi(arg[4]);
    } else if (m == "atom") {
      if (narg != 3) fail("Illegal write_dump_neighbor_debug command");
      single = 2;
      id = std::atol(arg[2]);
    } else if (m == "nva") {
      fail("write_dump_neighbor_debug nva: the device lists have no rows for non-Gauss virtual atoms");
    } else {
      fail("Illegal write_dump_neighbor_debug command");
    }
  }
  const Elements& el = s.element;
  const int* nc = el.ncell;
  const int A = s.atom.n;
  const double L[3] = {s.boxhi[0] - s.boxlo[0], s.boxhi[1] - s.boxlo[1], s.boxhi[2] - s.boxlo[2]};
  // the row of the requested point (atom rows first, then the Gauss cells of
  // each element in element order, as the reference's ilist orders them)
  long point = -1;
  if (single == 1) {
    if (ibasis != 0 || igcell < 0 || igcell > 7) return "";
    for (int e = 0; e < el.n; e++)
      if (el.tag[e] == id) point = A + 8L * e + igcell;
  } else if (single == 2) {
    for (int a = 0; a < A; a++)
      if (s.atom.tag[a] == id) point = a;
  }
  if (single && point < 0) return "";  // not owned here, as the reference returns
  const std::string path = step_name(arg[0], ntimestep, single ? "_proc_0" : "_proc_0.dat");
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) fail("Cannot open data file " + path);
  const long P = A + 8L * el.n;
  const long n = single ? nl.first[point + 1] - nl.first[point] + 1 : P;
  // the reference prints the rank in the TIMESTEP item (write_dump_neighbor_debug.cpp:377-378)
  std::fprintf(fp, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%ld\n", 0, n);
  box_header(fp, s);
  std::fprintf(fp, single ? "ITEM: ATOMS id type x y z tag iucell ibasis\n"
                          : "ITEM: ATOMS id type x y z tag numneigh igcell ibasis\n");
  auto site_pos = [&](int kind, int owner, int u, double* x) {
    if (kind == 0) {
      for (int d = 0; d < 3; d++) x[d] = s.atom.x[(size_t)owner * 3 + d];
    } else {
      double N[8];
      hex_shape(nc, u, N);
      for (int d = 0; d < 3; d++) x[d] = interp(el.nodex, owner, N, d);
    }
  };
  double x[3];
  long count = 1;
  if (!single) {
    for (long p = 0; p < P; p++) {
      const long nn = nl.first[p + 1] - nl.first[p];
      if (p < A) {
        site_pos(0, (int)p, 0, x);
        std::fprintf(fp, "%ld %d %g %g %g %d %ld %d %d\n", count++, s.atom.type[p], x[0], x[1], x[2], s.atom.tag[p],
                     nn, -1, -1);
      } else {
        const int e = (int)((p - A) / 8), g = (int)((p - A) % 8);
        site_pos(1, e, gauss_ucell(nc, g), x);
        std::fprintf(fp, "%ld %d %g %g %g %d %ld %d %d\n", count++, el.ctype[e], x[0], x[1], x[2], el.tag[e], nn, g,
                     0);
      }
    }
  } else {
    int iucell = -2, ib = -2;
    if (single == 1) {
      iucell = gauss_ucell(nc, igcell);
      ib = ibasis;
      site_pos(1, (int)((point - A) / 8), iucell, x);
    } else {
      site_pos(0, (int)point, 0, x);
    }
    std::fprintf(fp, "%ld 1 %g %g %g %ld %d %d\n", count++, x[0], x[1], x[2], id, iucell, ib);
    for (long q = nl.first[point]; q < nl.first[point + 1]; q++) {
      const int k = nl.kind[q], o = nl.owner[q], u = nl.ucell[q];
      site_pos(k, o, u, x);
      for (int d = 0; d < 3; d++) x[d] += nl.img[3 * (size_t)q + d] * L[d];
      if (k == 0)
        std::fprintf(fp, "%ld 2 %g %g %g %d %d %d\n", count++, x[0], x[1], x[2], s.atom.tag[o], -1, -1);
      else
        std::fprintf(fp, "%ld 3 %g %g %g %d %d %d\n", count++, x[0], x[1], x[2], el.tag[o], u, 0);
    }
  }
  std::fclose(fp);
  return path;
}

}  // namespace cac_host
