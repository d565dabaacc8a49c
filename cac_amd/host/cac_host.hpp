// cac_host.hpp -- host-side C++ layer over the C ABI (include/cac_hip.h).
//
// Mirrors the reference's plugin interfaces for the force + neighbor path so
// that a timestep loop written against them reads the same:
//   Pair           pair.h:92-121 (settings, coeff, init_style, init_one,
//                  compute(eflag, vflag), eng_vdwl, virial[6])
//   PairEAMAlloy   pair_eam_alloy.cpp:100-315 ("pair_coeff * * file.setfl El")
//   PairLJCut      pair_lj_cut.cpp:142-211 ("pair_style lj/cut rc",
//                  "pair_coeff 1 1 eps sigma")
//   Neighbor       neighbor.cpp:61-69 (skin, every, delay, check, one),
//                  decide() neighbor.cpp:1584-1604, build() 1704-1873
//   FixNVE         fix_nve.cpp:52-217 (initial_integrate, final_integrate)
//   Verlet         verlet.cpp:65-360 (setup, run)
//   read_data      read_data.cpp:338-375 with the section formats of
//                  atom_vec_atomic.cpp:26 and element_vec_cac.cpp:300-391
//
// Every object shares one Device, which owns the cac_hip_ctx: the model
// state lives in HBM, the host keeps the reference's AoS arrays and copies
// them in and out only when asked.  Errors from the C ABI throw cac_host::Error
// with the library's message, the analogue of error->all.
#ifndef CAC_HOST_HPP
#define CAC_HOST_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cac_hip.h"

namespace cac_host {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Atom (atom.h:36-39): x, v, f [n][3], type [n], tag [n]
struct Atoms {
  int n = 0;
  std::vector<double> x, v, f;
  std::vector<int> type, tag;
};

// Element (element.h:47-95), Hex with one basis atom: nodex/nodev/nodef
// [n][8][3], gaussf [n][8 Gauss cells][3], etype, ctype, tag [n]
struct Elements {
  int n = 0;
  int ncell[3] = {0, 0, 0};
  std::vector<double> nodex, nodev, nodef, gaussf;
  std::vector<int> etype, ctype, tag;
};

// One rank's model: units (0 lj, 1 metal), box, masses[1..ntypes]
struct System {
  int units = 1;
  double boxlo[3] = {0, 0, 0}, boxhi[3] = {0, 0, 0};
  std::vector<double> mass;  // mass[t - 1] of type t
  Atoms atom;
  Elements element;
};

// read_data subset: header counts and box, Masses, Element Types (Hex, apc 1,
// Gauss cells 2 2 2), Atoms, Velocities, Elements, Nodes, Node Velocities.
// Entries are placed by id, so sections may list them in any order.
void read_data(const std::string& path, System& sys);

// write_data (write_data.cpp:188-600) of a model held on the host: header
// (write_data.cpp:262-283), Masses (:290-294, "%d %g" as there), Element
// Types (element_vec_cac.cpp:1283-1292), Atoms (atom_vec_atomic.cpp:588-594),
// Velocities (atom_vec.cpp:101-106), Elements (element_vec_cac.cpp:1177-1182,
// the centre as the mean of the nodes), Nodes (:1203-1212) and Node
// Velocities (:1256-1264), in the model's order.  Coordinates and velocities
// print as %-1.16e (17 significant digits), so read_data reads every double
// back exactly.  velocities = false is the command's novelocity keyword.
void write_data_file(const std::string& path, const System& sys, long ntimestep, bool velocities = true);

class Device;
// the write_data command on the device model: Domain::pbc on the device, as
// WriteData::command runs it (write_data.cpp:120-131), then the state
// downloaded into the System and written.  The lists are stale afterwards:
// the next run's setup rebuilds them, as the reference's does.
void write_data(const std::string& path, Device& dev, long ntimestep, bool velocities = true);

// The context and the uploaded model.  precision: CAC_HIP_FP64 or CAC_HIP_MIXED.
class Device {
 public:
  Device(System& sys, int device = 0);
  ~Device();
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  cac_hip_ctx* ctx() const { return ctx_; }
  System& system() { return sys_; }
  void check(int rc) const;            // throw Error on a non-zero return
  void upload();                       // model arrays host -> device (invalidates lists)
  void sync_positions();               // atom->x, element->nodex host -> device
  void download_state();               // x, v (and nodal x, v) device -> host
  void pbc();                          // Domain::pbc on the device model (lists stale after)
  void download_forces();              // f, nodef, gaussf device -> host

 private:
  System& sys_;
  cac_hip_ctx* ctx_ = nullptr;
};

class Pair {
 public:
  explicit Pair(Device& dev) : dev_(dev) {}
  virtual ~Pair() = default;
  virtual void settings(int narg, char** arg) = 0;
  virtual void coeff(int narg, char** arg) = 0;
  virtual void init_style() {}           // the neighbor request is implicit: full lists on device
  virtual double init_one(int i, int j) = 0;
  // Pair::compute + the reverse comm + Element::compute_nodef, on device;
  // eflag / vflag as force.h:11-19; fills eng_vdwl and virial[6], and with
  // ENERGY_ATOM / VIRIAL_ATOM (or VIRIAL_CENTROID) the per-point tallies
  virtual void compute(int eflag, int vflag);
  double eng_vdwl = 0.0;
  double virial[6] = {0, 0, 0, 0, 0, 0};
  // pair.h:13-17: eatom [natoms], enode [nelements][8], vatom [natoms][6], vnode [nelements][8][6]
  std::vector<double> eatom, enode, vatom, vnode;

 protected:
  Device& dev_;
};

class PairEAMAlloyHIP : public Pair {
 public:
  using Pair::Pair;
  void settings(int narg, char** arg) override;  // no arguments
  void coeff(int narg, char** arg) override;     // * * file.setfl Element
  double init_one(int i, int j) override;
  double cutmax = 0.0;
};

class PairLJCutHIP : public Pair {
 public:
  using Pair::Pair;
  void settings(int narg, char** arg) override;  // cut_global
  void coeff(int narg, char** arg) override;     // 1 1 epsilon sigma [cut]
  double init_one(int i, int j) override;
  double cut_global = 0.0, epsilon = 1.0, sigma = 1.0, cut = 0.0;
};

class NeighborHIP {
 public:
  explicit NeighborHIP(Device& dev) : dev_(dev) {}
  // "neighbor skin bin" and "neigh_modify every N delay N check yes one N"
  double skin = 2.0;
  int every = 1, delay = 10, check = 1, oneatom = 2000;
  void init();      // hand the settings to the device
  int decide();     // 1 when a rebuild is due (Neighbor::decide)
  void build();     // Domain::pbc + ghost shell + Neighbor::build, on device
  int ncalls() const;
  int ndanger() const;

 private:
  Device& dev_;
};

class FixNVEHIP {
 public:
  FixNVEHIP(Device& dev, double dt) : dev_(dev), dt_(dt) {}
  void initial_integrate(int vflag);
  void final_integrate();
  void reset_dt(double dt) { dt_ = dt; }

 private:
  Device& dev_;
  double dt_;
};

// Thermo row (compute_pe / compute_temp / compute_pressure, thermo.cpp:1929-1936)
struct ThermoRow {
  long step;
  double pe, ke, etotal, press, temp;
};

class VerletHIP {
 public:
  VerletHIP(Device& dev, Pair& pair, NeighborHIP& neigh, FixNVEHIP& fix)
      : dev_(dev), pair_(pair), neigh_(neigh), fix_(fix) {}
  // Verlet::setup: neighbor build + first force with energy and virial
  void setup();
  // Verlet::run step by step through the plugin objects, as verlet.cpp:218-360
  // orders them; thermo every thermo_every steps and on the last
  std::vector<ThermoRow> run(int nsteps, int thermo_every);
  // the same loop resident on the device (cac_hip_run): no host round trip per step
  std::vector<ThermoRow> run_resident(int nsteps, double dt, int thermo_every);
  ThermoRow thermo(long step);
  long ntimestep = 0;

 private:
  Device& dev_;
  Pair& pair_;
  NeighborHIP& neigh_;
  FixNVEHIP& fix_;
};

// ------------------------------------------------------------------ output taps (cac_dump.cpp)

// "dump ID all custom N file [format F] [average yes|no] [wrap yes|no]
// [precision high|normal] col ..." in the LAMMPS text format
// (dump_custom.cpp:30-121, 1260-1420, 1672-1719).  Formats lammps/full/map
// (default), lammps/surface, lammps/gauss/point, lammps/node, lammps/center;
// columns id type x y z vx vy vz fx fy fz mass proc procp1 eid etype ibasis
// igauss inode iucell.  Real atoms come first, then each element's virtual
// atoms, whose velocity and force columns are interpolated from nodev / nodef.
// A file name with '*' gives one file per snapshot; otherwise snapshots are
// appended to one file.
class DumpCustom {
 public:
  DumpCustom(int narg, char** arg);
  void write(const System& s, long ntimestep);
  long count(const System& s) const;  // lines per snapshot
  std::string id, file, columns;
  int nevery = 1;

 private:
  enum Style { FULLMAP, SURFACE, GAUSSPOINT, NODE, CENTER };
  enum Col { ID, TYPE, X, Y, Z, VX, VY, VZ, FX, FY, FZ, MASS, PROC, PROCP1, EID, ETYPE, IBASIS, IGAUSS, INODE, IUCELL };
  void row(std::vector<double>& buf, const System& s, int kind, int i, int ibasis, int iucell, int igcell, int inode,
           long vid, const double* x) const;
  Style style = FULLMAP;
  std::vector<Col> cols;
  std::vector<bool> col_int;
  bool wrap = false, precision_high = false, opened_ = false;
};

// The device's full lists by point (atoms, then 8 Gauss cells per element):
// entries first[p] .. first[p+1]-1 are (kind 0 atom / 1 virtual atom, owner,
// unit cell, image shift); the neighbor sits at owner's position + img * L.
struct NeighborRows {
  std::vector<long> first;
  std::vector<int> kind, owner, ucell, img;
};
NeighborRows export_neighbor_rows(Device& dev);

// "write_dump_neighbor_debug file [gauss id ibasis igcell | atom id]"
// (write_dump_neighbor_debug.cpp:41-409): every row with its list length, or
// one row's neighbors with their positions, tags and unit cells.  The device
// keeps full rows, so numneigh and a row's entries are those of the full
// list: the reference's half row of a point is a subset.  Returns the path
// written ("" when the point does not exist, as the reference writes nothing).
std::string write_dump_neighbor_debug(const System& s, const NeighborRows& nl, int narg, char** arg, long ntimestep);

}  // namespace cac_host

#endif
