// cac_ctx.hpp -- device-resident state of one CAC model and the launchers the
// C ABI (api.hip) sequences.  One process drives one MI355X; every launch goes
// on the context's stream.
//
// HBM layout (SoA, FP64 unless noted):
//   points  P = A + 8E: atoms [0, A), then the 8 nodes of element e at A + 8e + k
//           (Hex node order).  x/v/f are [3][P] planes; f holds atom forces and
//           assembled nodal forces (nodef).  gf [3][8E] is the Gauss-cell force at
//           each node (2x2x2 Gauss cells sit on the nodes).
//   sites   S = P + needed virtual atoms, sorted by cell.  sx [3][S] is the site
//           position wrapped into the box; sdesc names the source (point index >= 0,
//           or -(1 + e*nucell + u) for unit cell u of element e); simg is the box
//           shift applied when wrapping; sfp is F'(rho) at the site (EAM).
//   rows    R = P points in cell order; nlist is column-major [cap][R] of
//           (site | image << 27), image = (tx+1) + 3(ty+1) + 9(tz+1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace cac {

constexpr int NPE = 8;
constexpr int IMG_SHIFT = 27;
constexpr unsigned SITE_MASK = (1u << IMG_SHIFT) - 1u;  // list entry = site | image << 27, read as unsigned
constexpr int MAXT = 8;

struct HipError {
  hipError_t code;
  std::string where;
};

#define CAC_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) throw ::cac::HipError{e_, std::string(#call)};      \
  } while (0)

// Device buffer that grows (never shrinks) and frees on destruction.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  void reserve(size_t count) {
    if (count <= n && p) return;
    if (p) CAC_HIP(hipFree(p));
    p = nullptr;
    size_t c = count ? count : 1;
    CAC_HIP(hipMalloc(&p, c * sizeof(T)));
    n = c;
  }
  T* get() const { return p; }
};

struct Timer {
  double sum_us = 0.0;
  long count = 0;
};

struct ErrorFlags {  // device-written status words
  int nan;           // non-numeric coordinate seen
  int deform;        // element deformed > 2x
  int overflow;      // neighbor row over capacity
  int maxcount;      // largest row seen in the last build
  int rebuild;       // check_distance result
  int pad[3];
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // units (update.cpp:112-184)
  int units = 1;
  double boltz = 8.617343e-5, mvv2e = 1.0364269e-4, ftm2v = 1.0 / 1.0364269e-4, nktv2p = 1.6021765e6;

  // box
  bool box_set = false;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, L[3] = {0, 0, 0};

  // types
  int ntypes = 0;
  double mass[MAXT + 1] = {0};

  // element type (one Hex type, apc 1, 2x2x2 Gauss cells at the nodes)
  bool etype_set = false;
  int ncell[3] = {0, 0, 0}, nucell = 0, subsplit = 1;
  double nodal_weight = 0.0;
  std::vector<double> shape_h;  // [nucell][8]
  std::vector<int> u2g_h;       // [nucell]
  int g2u[8], g2n[8], n2g[8];
  double wsa[8][8], wsa_lumped[8][8];  // [g][node]
  double minv[8][8], mmat[8][8];
  DBuf<double> shape;  // device copy [nucell][8]
  DBuf<int> u2g;

  // host staging of the uploaded model (reference AoS layouts)
  std::vector<double> st_ax, st_av, st_nx, st_nv;
  std::vector<int> st_at, st_ct;

  // model sizes
  int A = 0, E = 0, P = 0;
  bool model_set = false;

  // points
  DBuf<double> x, v, f, gf, xhold, ex, initsize, rho;
  DBuf<int> ptype;

  // sites / cells / rows / lists
  int S = 0, nvirt = 0;
  int nc[3] = {0, 0, 0};
  long ncells = 0;
  double cinv[3] = {0, 0, 0};
  DBuf<double> sx, sfp, tsx;
  DBuf<float> sfp32;
  DBuf<int> sdesc, simg, tdesc, timg, tcell, perm, point_site, row_point, row_site, ecount, eoffset, scan_tmp;
  DBuf<int> cell_count, cell_start, cell_cursor, occ, nearA, nearB;
  DBuf<int> nnum, nlist;
  int cap = 0;
  long entries = 0;
  DBuf<int> isflag, rowidx;

  // neighbor settings (neighbor.cpp:61-69, update.cpp:137,184)
  double skin = 2.0, cutforce = 0.0, cutneigh = 0.0, cutneighsq = 0.0, triggersq = 0.0;
  int every = 1, delay = 10, check = 1, one = 2000;
  int ago = 0, ncalls = 0, ndanger = 0;
  bool built = false;
  long ntimestep = 0;

  // pair
  int pairstyle = 0;  // 1 lj/cut, 2 eam/alloy
  double lj1 = 0, lj2 = 0, lj3 = 0, lj4 = 0, ljcutsq = 0;
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, rdr = 0, rdrho = 0, rhomax = 0, cutmax = 0, cutforcesq = 0;
  DBuf<double> frho_spl, rhor_spl, z2r_spl;  // [(n+1)*7]
  DBuf<float> frho_spl32, rhor_spl32, z2r_spl32;
  int precision = 0;     // 0 FP64, 1 mixed
  int nodef_style = 0;   // 0 consistent, 1 lumped, 2 node/only

  // energy / virial
  double eng_vdwl = 0.0, virial[6] = {0, 0, 0, 0, 0, 0};
  DBuf<double> erow;      // [7][R]: energy then 6 virial components per row
  DBuf<double> partials;  // block partial sums
  double* h_red = nullptr;  // pinned host scratch for reductions
  ErrorFlags* h_flags = nullptr;  // pinned mirror of the device flags
  DBuf<ErrorFlags> d_flags;

  // timing
  bool timing = false;
  std::map<std::string, Timer> timers;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  ~Ctx();
};

// ---------------------------------------------------------------- launchers
// neigh.hip
void launch_pbc(Ctx& c);
void check_deformation(Ctx& c, bool first);
void build_neighbor(Ctx& c);  // pbc + sites + cells + full lists
int check_distance(Ctx& c);   // returns 1 when any point moved more than skin/2
void export_lists(Ctx& c, std::vector<int>& row, std::vector<int>& kind, std::vector<int>& owner,
                  std::vector<int>& ucell, std::vector<int>& img);
void scan_exclusive(Ctx& c, const int* in, int* out, int n, int* total);

// pair.hip
void site_refresh(Ctx& c);
void pair_compute(Ctx& c, int eflag, int vflag);

// integrate.hip
void nve_initial(Ctx& c, double dt);
void nve_final(Ctx& c, double dt);
void update_centers(Ctx& c);
void nodal_assembly(Ctx& c);
void thermo(Ctx& c, double* ke, double* temp);
void reduce_rows(Ctx& c, const double* data, int n, int ncomp, double* out);

// timing helpers (api.hip)
struct ScopedTimer {
  Ctx& c;
  const char* name;
  ScopedTimer(Ctx& cc, const char* nm);
  ~ScopedTimer();
};

inline int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

}  // namespace cac
