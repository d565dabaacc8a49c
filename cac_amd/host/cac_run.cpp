// cac_run.cpp -- command-line driver of the host layer: read a CAC data file,
// set up a pair style, run velocity-Verlet steps on the device and print
// thermo rows the way the reference's thermo output lists them
// ("Step PotEng KinEng TotEng Press").
//
//   cac_run DATA (--eam SETFL | --lj EPS SIGMA CUT) [--units metal|lj]
//           [--skin S] [--every N] [--delay N] [--check yes|no] [--dt DT]
//           [--steps N] [--thermo N] [--mixed] [--loop host|device]
//           [--dump "ID all custom N file ..."]... [--neighbor-debug "file [atom ID | gauss ID IB IG]"]
//           [--write-data FILE [--then M FILE2]]
//   cac_run DATA --parse-only [--dump ...] [--write-data FILE]   counts and checksums of the data file
//
// --dump takes the arguments of the reference's dump command and writes a
// snapshot at step 0 and every N steps; --neighbor-debug those of
// write_dump_neighbor_debug, written after the first neighbor build.
// --write-data is a write_data command after the run; --then M FILE2 a second
// run command of M steps (with its own setup) and a write_data after it.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "cac_host.hpp"

using namespace cac_host;

static int usage() {
  std::fprintf(stderr,
               "usage: cac_run DATA (--eam SETFL | --lj EPS SIGMA CUT) [--units metal|lj] [--skin S]\n"
               "               [--every N] [--delay N] [--check yes|no] [--dt DT] [--steps N] [--thermo N]\n"
               "               [--mixed] [--loop host|device] [--dump \"ID all custom N file ...\"]...\n"
               "               [--neighbor-debug \"file [atom ID | gauss ID IBASIS IGCELL]\"]\n"
               "               [--write-data FILE [--then M FILE2]]\n"
               "       cac_run DATA --parse-only [--dump ...] [--write-data FILE]\n");
  return 2;
}

// split a command's arguments on whitespace, as the input parser hands them on
struct Args {
  std::vector<std::string> w;
  std::vector<char*> p;
  explicit Args(const std::string& line) {
    std::istringstream in(line);
    for (std::string t; in >> t;) w.push_back(t);
    for (auto& t : w) p.push_back(&t[0]);
  }
  int n() { return (int)p.size(); }
  char** v() { return p.data(); }
};

int main(int argc, char** argv) {
  if (argc < 3) return usage();
  std::vector<std::string> dump_args;
  std::string ndebug_args;
  std::string data = argv[1], eam, units = "metal", loop = "device", wdata, wdata2;
  long then_steps = -1;
  double eps = 1, sigma = 1, cut = 2.5, skin = -1, dt = -1;
  int every = 1, delay = 10, check = 1, steps = 100, thermo = 10;
  bool lj = false, mixed = false, parse_only = false;
  for (int i = 2; i < argc; i++) {
    std::string a = argv[i];
    auto need = [&](int k) {
      if (i + k >= argc) std::exit(usage());
    };
    if (a == "--eam") { need(1); eam = argv[++i]; }
    else if (a == "--lj") { need(3); lj = true; eps = std::atof(argv[++i]); sigma = std::atof(argv[++i]); cut = std::atof(argv[++i]); }
    else if (a == "--units") { need(1); units = argv[++i]; }
    else if (a == "--skin") { need(1); skin = std::atof(argv[++i]); }
    else if (a == "--every") { need(1); every = std::atoi(argv[++i]); }
    else if (a == "--delay") { need(1); delay = std::atoi(argv[++i]); }
    else if (a == "--check") { need(1); check = std::strcmp(argv[++i], "no") != 0; }
    else if (a == "--dt") { need(1); dt = std::atof(argv[++i]); }
    else if (a == "--steps") { need(1); steps = std::atoi(argv[++i]); }
    else if (a == "--thermo") { need(1); thermo = std::atoi(argv[++i]); }
    else if (a == "--mixed") mixed = true;
    else if (a == "--loop") { need(1); loop = argv[++i]; }
    else if (a == "--parse-only") parse_only = true;
    else if (a == "--dump") { need(1); dump_args.push_back(argv[++i]); }
    else if (a == "--neighbor-debug") { need(1); ndebug_args = argv[++i]; }
    else if (a == "--write-data") { need(1); wdata = argv[++i]; }
    else if (a == "--then") { need(2); then_steps = std::atol(argv[++i]); wdata2 = argv[++i]; }
    else return usage();
  }
  try {
    System sys;
    sys.units = units == "lj" ? CAC_HIP_UNITS_LJ : CAC_HIP_UNITS_METAL;
    read_data(data, sys);
    std::vector<std::unique_ptr<DumpCustom>> dumps;
    for (auto& d : dump_args) {
      Args a(d);
      dumps.emplace_back(new DumpCustom(a.n(), a.v()));
    }
    if (then_steps >= 0 && wdata.empty()) return usage();
    if (parse_only) {
      for (auto& d : dumps) d->write(sys, 0);
      if (!wdata.empty()) write_data_file(wdata, sys, 0);
      double sx = 0, sv = 0, sn = 0, snv = 0;
      for (double q : sys.atom.x) sx += q;
      for (double q : sys.atom.v) sv += q;
      for (double q : sys.element.nodex) sn += q;
      for (double q : sys.element.nodev) snv += q;
      std::printf("atoms %d elements %d types %zu ncell %d %d %d\n", sys.atom.n, sys.element.n, sys.mass.size(),
                  sys.element.ncell[0], sys.element.ncell[1], sys.element.ncell[2]);
      std::printf("box %.17g %.17g %.17g %.17g %.17g %.17g\n", sys.boxlo[0], sys.boxhi[0], sys.boxlo[1],
                  sys.boxhi[1], sys.boxlo[2], sys.boxhi[2]);
      std::printf("sums %.17g %.17g %.17g %.17g\n", sx, sv, sn, snv);
      return 0;
    }
    if (eam.empty() && !lj) return usage();
    if (skin < 0) skin = sys.units == CAC_HIP_UNITS_METAL ? 2.0 : 0.3;  // update.cpp:112-184 defaults
    if (dt < 0) dt = sys.units == CAC_HIP_UNITS_METAL ? 0.001 : 0.005;
    Device dev(sys);
    PairEAMAlloyHIP pe(dev);
    PairLJCutHIP pl(dev);
    Pair* pair = &pe;
    if (lj) {
      char c0[32], e0[32], s0[32];
      std::snprintf(c0, sizeof c0, "%.17g", cut);
      std::snprintf(e0, sizeof e0, "%.17g", eps);
      std::snprintf(s0, sizeof s0, "%.17g", sigma);
      char one[] = "1";
      char* sa[] = {c0};
      char* ca[] = {one, one, e0, s0};
      pl.settings(1, sa);
      pl.coeff(4, ca);
      pair = &pl;
    } else {
      char star[] = "*", el[] = "X";
      std::string f = eam;
      char* ca[] = {star, star, &f[0], el};
      pe.settings(0, nullptr);
      pe.coeff(4, ca);
    }
    dev.check(cac_hip_set_precision(dev.ctx(), mixed ? CAC_HIP_MIXED : CAC_HIP_FP64));
    NeighborHIP neigh(dev);
    neigh.skin = skin;
    neigh.every = every;
    neigh.delay = delay;
    neigh.check = check;
    FixNVEHIP fix(dev, dt);
    VerletHIP verlet(dev, *pair, neigh, fix);
    verlet.setup();
    if (!ndebug_args.empty()) {
      Args a(ndebug_args);
      dev.download_state();
      const std::string f = write_dump_neighbor_debug(sys, export_neighbor_rows(dev), a.n(), a.v(), 0);
      if (!f.empty()) std::printf("neighbor debug written to %s\n", f.c_str());
    }
    auto snapshot = [&](long step) {
      bool due = false;
      for (auto& d : dumps) due = due || step % d->nevery == 0;
      if (!due) return;
      dev.download_state();
      dev.download_forces();
      for (auto& d : dumps)
        if (step % d->nevery == 0) d->write(sys, step);
    };
    snapshot(0);
    // run in chunks that end on every dump step; thermo rows keep their own
    // cadence, the rows the chunk ends add are dropped
    long chunk = steps;
    for (auto& d : dumps) chunk = std::min<long>(chunk, d->nevery);
    if (chunk <= 0) chunk = steps;
    std::vector<ThermoRow> rows;
    long done = 0;
    while (done < steps || rows.empty()) {
      const int n = (int)std::min<long>(chunk, steps - done);
      auto part = loop == "host" ? verlet.run(n, thermo) : verlet.run_resident(n, dt, thermo);
      for (size_t k = rows.empty() ? 0 : 1; k < part.size(); k++) {
        const long st = part[k].step;
        if (k == 0 || st == steps || (thermo > 0 && st % thermo == 0)) rows.push_back(part[k]);
      }
      if (n == 0) break;
      done += n;
      snapshot(done);
    }
    std::printf("Step PotEng KinEng TotEng Press\n");
    for (const auto& r : rows) std::printf("%ld %.15g %.15g %.15g %.15g\n", r.step, r.pe, r.ke, r.etotal, r.press);
    std::printf("Neighbor list builds = %d\nDangerous builds = %d\n", neigh.ncalls(), neigh.ndanger());
    if (!wdata.empty()) write_data(wdata, dev, verlet.ntimestep);
    if (then_steps >= 0) {
      // the next run command: Verlet::setup (a neighbor build) and its steps
      verlet.setup();
      if (then_steps > 0) {
        if (loop == "host") verlet.run((int)then_steps, 0);
        else verlet.run_resident((int)then_steps, dt, 0);
      }
      write_data(wdata2, dev, verlet.ntimestep);
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  return 0;
}
