// cac_ctx.hpp -- device-resident state of one CAC model and the launchers the
// C ABI (api.hip) sequences.  One process drives one MI355X; every launch goes
// on the context's stream.
//
// HBM layout (SoA, FP64 unless noted):
//   points  P = A + 8E: atoms [0, A), then the 8 nodes of element e at A + 8e + k
//           (Hex node order).  x/v/f are [3][P] planes; f holds atom forces and
//           assembled nodal forces (nodef).  gf [3][8E] is the Gauss-cell force at
//           each node (2x2x2 Gauss cells sit on the nodes).
//   sites   S = P + the virtual atoms some outer list references, in cell order.
//           sdesc names the source (point index >= 0, or -(1 + e*nucell + u) for
//           unit cell u of element e); simg is the box shift applied when
//           wrapping; scell the site's neighbor cell.  Full position records
//           (site, FP64; site32f, FP32 cell-relative) are written on prune steps.
//   live    the sites some inner list reads, compacted after each prune into
//           dense arrays in cell order (lsdesc/lsimg/lscell).  site_live (FP64)
//           or site32 (mixed, relative to the site's cell corner) is a record
//           (x, y, z, F'(rho)), so one pair needs one gather.
//   rows    R = P points in Morton order of their cell; nlist is row-major
//           [R][cap] of unsigned (site | image << 27), image = (tx+1) + 3(ty+1)
//           + 9(tz+1): every site image within rc + skin (the reference's list).
//           ilist [R][icap] is the inner list the force passes walk: the entries
//           within rc + delta, pruned on device and rebuilt whenever a point has
//           moved delta/2 since, indexing the live records (FP64: live site |
//           image << 27; mixed: live site | cell offset << 25).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <utility>
#include <string>
#include <vector>

namespace cac {

constexpr int NPE = 8;
// inner-list rows are padded to a multiple of ROW_PAD entries with entries
// naming a dummy record beyond every cutoff; the pair passes read whole batches
constexpr int ROW_PAD = 16;
constexpr int IMG_SHIFT = 27;
constexpr unsigned SITE_MASK = (1u << IMG_SHIFT) - 1u;  // list entry = site | image << 27, read as unsigned

// a site's neighbor cell is kept as packed coordinates ix | iy << 10 | iz << 21;
// the top bit stays clear (the mixed path's full records set it for a site
// that is a point, POINT_SITE_BIT)
constexpr int CELL_BITS_X = 10, CELL_BITS_Y = 11, CELL_BITS_Z = 10;
constexpr unsigned POINT_SITE_BIT = 1u << 31;
__host__ __device__ inline int cell_pack(int ix, int iy, int iz) {
  return ix | (iy << CELL_BITS_X) | (iz << (CELL_BITS_X + CELL_BITS_Y));
}
__host__ __device__ inline int cell_ix(int p) { return p & ((1 << CELL_BITS_X) - 1); }
__host__ __device__ inline int cell_iy(int p) { return (p >> CELL_BITS_X) & ((1 << CELL_BITS_Y) - 1); }
__host__ __device__ inline int cell_iz(int p) { return (p >> (CELL_BITS_X + CELL_BITS_Y)) & ((1 << CELL_BITS_Z) - 1); }

// Workgroup b of a launch runs on XCD b % 8.  xcd_block renumbers the grid so
// that each XCD's workgroups take one contiguous range of logical blocks: data
// that neighbouring blocks share then stays in one XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int grid) {
  const int q = grid / 8, rem = grid % 8, x = b % 8;
  return x * q + (x < rem ? x : rem) + b / 8;
}

// Trilinear Hex shape functions of unit cell u (i-major: u = (i ny + j) nz + k)
// at natural coordinates (h_x i - 1, h_y j - 1, h_z k - 1), h = 2 / (n - 1),
// formed with the same operations, in the same order and unfused, as the
// element set-up forms its shape table (element.cpp:2989-2997), so the values
// are bitwise those of the table.
struct UcellGrid {
  double h[3];
  int ny, nz, nucell;
  double inv_nucell;      // 1 / nucell
  float inv_nyz, inv_nz;  // 1 / (ny nz), 1 / nz
};

// a / b for 0 <= a < 2^31, b > 0, from an estimate off by at most one
__device__ __forceinline__ int div_fix(int a, int b, int est) {
  est -= est * b > a ? 1 : 0;
  est += (est + 1) * b <= a ? 1 : 0;
  return est;
}

// element e and unit cell u of the interpolated-atom descriptor q = e nucell + u
__device__ __forceinline__ void split_ucell(const UcellGrid& ug, int q, int& e, int& u) {
  e = div_fix(q, ug.nucell, (int)((double)q * ug.inv_nucell));
  u = q - e * ug.nucell;
}

__device__ __forceinline__ void hex_shape_at(const UcellGrid& ug, int u, double* N) {
#pragma clang fp contract(off)
  const int nyz = ug.ny * ug.nz;
  const int i = div_fix(u, nyz, (int)((float)u * ug.inv_nyz));
  const int r = u - i * nyz;
  const int j = div_fix(r, ug.nz, (int)((float)r * ug.inv_nz));
  const int k = r - j * ug.nz;
  const double px = ug.h[0] * i - 1.0, py = ug.h[1] * j - 1.0, pz = ug.h[2] * k - 1.0;
  N[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
  N[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
  N[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
  N[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
  N[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
  N[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
  N[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
  N[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
}
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
  // grows only, with a quarter of headroom, so that sizes that drift from
  // one neighbor build to the next do not reallocate every time
  void reserve(size_t count) {
    if (count <= n && p) return;
    if (p) CAC_HIP(hipFree(p));
    p = nullptr;
    size_t c = count ? count + count / 4 : 1;
    CAC_HIP(hipMalloc(&p, c * sizeof(T)));
    n = c;
  }
  T* get() const { return p; }
  void swap(DBuf& o) {
    T* tp = p;
    p = o.p;
    o.p = tp;
    size_t tn = n;
    n = o.n;
    o.n = tn;
  }
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
  int rebuild;       // check_distance result (outer list)
  int rebuild_in;    // inner list no longer covers rc
  int rows;          // the sites that are points did not number P rows (an internal fault)
  int pad;
};

// Gate of the speculative no-rebuild step: the host queues a whole step
// before it knows whether the displacement check asked for a rebuild; the
// check writes the step number into steps[0] (rebuild) / steps[1] (re-prune
// the inner list), and gated kernels of that step return at once on a
// rebuild.  step < 0: not gated.
struct Gate {
  const int* steps = nullptr;
  int step = -1;
  __device__ bool skip() const { return step >= 0 && steps[0] == step; }
  __device__ bool prune() const { return step < 0 || (steps[1] == step && steps[0] != step); }
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
  DBuf<double4> site;
  DBuf<float4> site32;  // mixed path: (x, y, z) - cell corner, F'(rho)
  DBuf<int> scell, scell2;  // cell of each site (the frame of site32)
  DBuf<int> sdesc, simg, sdesc2, simg2, used, newidx;
  DBuf<int> tcell, row_point, row_site, ecount, eoffset, scan_tmp;
  DBuf<int> cell_start, occ, nearA, nearB, sort_keys;
  DBuf<unsigned> tval, sort_vals;  // candidate site as (descriptor index) << 5 | wrap code, unsorted / sorted
  DBuf<unsigned long long> row_key;          // Morton keys of the cells (and their sorted copy)
  DBuf<int> row_idx, row_tmp;
  DBuf<int> cell_morton, morton_rows;  // cells in Morton order (per grid shape); rows per cell and their starts
  int morton_nc[3] = {0, 0, 0};        // the grid shape cell_morton was sorted for
  DBuf<unsigned char> sort_tmp;
  DBuf<int> nnum, nlist;
  DBuf<int> nearrow;  // per row: a candidate lies within the near-cutoff band
  int cap = 0;
  DBuf<int> innum, ilist;
  int icap = 0;
  double inner_skin = 0.0;       // delta in use
  double inner_skin_set = -1.0;  // requested delta; < 0 picks 0.3 skin
  double inner_cutsq = 0.0;
  int inner_mode = -1;       // precision the inner list was built for (-1: stale)
  long inner_builds = 0;
  int inner_due = 0;         // set by cac_hip_neigh_decide
  bool x_dirty = false;      // points moved since the sites were last refreshed
  DBuf<double> xhold_in;
  long entries = 0;
  DBuf<int> isflag, rowidx;

  // neighbor settings (neighbor.cpp:61-69, update.cpp:137,184)
  double skin = 2.0, cutforce = 0.0, cutneigh = 0.0, cutneighsq = 0.0, triggersq = 0.0;
  int every = 1, delay = 10, check = 1, one = 2000;
  int ago = 0, ncalls = 0, ndanger = 0;
  bool built = false;
  long ntimestep = 0;
  long begun_step = -1;  // a step whose initial kick, drift and checks already ran (fused)

  // pair
  int pairstyle = 0;  // 1 lj/cut, 2 eam/alloy
  double lj1 = 0, lj2 = 0, lj3 = 0, lj4 = 0, ljcutsq = 0;
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, rdr = 0, rdrho = 0, rhomax = 0, cutmax = 0, cutforcesq = 0;
  DBuf<double> frho_spl, rhor_spl, z2r_spl;  // [(n+1)*7]
  DBuf<float> tab32r;  // mixed density pass, per knot: rhor c3..c6 (FP32)
  DBuf<float> tab32f;  // mixed force pass: planes A, B, C of nr + 2 float4 (cubics in the knot fraction)
  DBuf<double> tab64;  // FP64 passes, per knot: rhor c3..c6 and z2r c3..c6 (the reference's coefficients)
  int precision = 0;     // 0 FP64, 1 mixed
  int tpr = 8;           // lanes per neighbor row in the density pass
  int tpr_force = 4;     // lanes per neighbor row in the force passes (EAM, LJ)
  int grid_cap = 4096;   // max workgroups of the persistent pair kernels
  int pair_shape = 2;    // EAM pass shape: 0/1/2/3 = (512|1024 threads) x (4|8 entries in flight)
  int dens_nt = 512;     // workgroup size of the EAM density pass (512 or 1024)
  int ncu = 256;         // compute units of the device
  std::map<std::pair<const void*, size_t>, int> resident;  // workgroups per CU of a kernel at an LDS size
  int nodef_style = 0;   // 0 consistent, 1 lumped, 2 node/only

  // energy / virial
  double eng_vdwl = 0.0, virial[6] = {0, 0, 0, 0, 0, 0};
  DBuf<double> erow;      // [7][R]: energy then 6 virial components per row
  DBuf<double> peratom;   // [7][P]: per-point energy and virial of the last pass that asked for them
  int peratom_flags = 0;  // 1: per-point energies, 2: per-point virials, of the last pair passes
  DBuf<double> partials;  // block partial sums
  double* h_red = nullptr;  // pinned host scratch for reductions
  ErrorFlags* h_flags = nullptr;  // pinned mirror of the device flags
  DBuf<int> d_steps;              // [rebuild step, prune step] written by the displacement check
  // live sites: the sites the inner lists read, compacted after every prune
  // into dense arrays in site (cell) order; the inner lists and the rows index them
  DBuf<int> live_idx, lsdesc, lsimg, lscell, row_live, d_nlive;
  DBuf<unsigned char> live_flag;  // one byte per site: the 4 MB of flags of a large model stay in L2
  size_t live_flag_clean = 0;      // leading flags known to be zero (the compaction clears what it reads)
  unsigned char* live_flag_at = nullptr;  // allocation those flags were cleared in
  DBuf<double4> site_live;        // FP64 records (x, y, z, F') of the live sites
  DBuf<float4> site32f;           // mixed path: FP32 cell-relative records of every site (for the prune)
  volatile int* h_steps = nullptr;  // host-mapped copy of the same, written by the check kernel
  int* h_steps_dev = nullptr;       // device address of h_steps
  // the check's completion: its kernels count their workgroups on per-XCD
  // device counters and a global one, and the last writes the step number
  // into h_steps[2]
  DBuf<unsigned> d_check_count;     // 9 counters, 64 words apart
  // host-mapped words through which a one-workgroup kernel hands small
  // device values (counts, status words) to the host: [0] sequence, [1..] values
  volatile int* h_pub = nullptr;
  int* h_pub_dev = nullptr;
  int pub_seq = 0;
  unsigned check_xcd[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // running targets of the XCD counters
  unsigned check_all = 0;                            // and of the global one
  hipEvent_t ev_seg = nullptr;      // recorded when the model is ready for segmented pair passes
  double* h_stage = nullptr;        // pinned staging of the per-step host handover
  size_t h_stage_n = 0;
  DBuf<ErrorFlags> d_flags;
  double* scratch = nullptr;  // device address of the allocation-free scratch array on this device

  // timing: an event pair on the context stream around each timed section,
  // read back only when the times are asked for (or the pool is full), so a
  // timed run is not synchronized section by section
  int timing = 0;  // 0 off, 1 every section, 2 the pair force pass only
  std::map<std::string, Timer> timers;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct PendingTime {
    const char* name;
    hipEvent_t start, stop;
  };
  std::vector<PendingTime> ev_pending;

  ~Ctx();
};

// ---------------------------------------------------------------- launchers
// neigh.hip
void launch_pbc(Ctx& c);
void wrap_model(Ctx& c);  // centres + pbc now, with the non-finite check
void check_deformation(Ctx& c, bool first);
void build_neighbor(Ctx& c);  // pbc + sites + cells + full lists
int check_distance(Ctx& c, bool outer, int* inner);  // 1 when any point moved more than skin/2
void build_inner(Ctx& c);  // prune the outer list to rc + delta and compact the live sites
// api.hip: one velocity-Verlet step in two halves (queue, then resolve the check)
bool step_issue(Ctx& c, double dt, int ev);
// fuse_next: the caller issues the next step right after this one and reads
// no full-step velocities in between (the closing kick is fused with the next
// step's opening kick, drift and checks)
void step_resolve(Ctx& c, double dt, int ev, bool fuse_next = false);
// step_resolve = step_prepare (wait for the check, rebuild or re-prune), the
// pair passes, step_close (nodal forces and the closing kick)
void step_prepare(Ctx& c);
void wait_check(Ctx& c);  // until the current step's displacement check has published its verdict
// small device values read back through the host-mapped words, in stream
// order (no copy engine, no stream synchronization): n <= 8 ints, or the sum
// of two ints
void read_device_ints(Ctx& c, const int* src, int n, int* dst);
int read_device_sum(Ctx& c, const int* a, const int* b);
void step_close(Ctx& c, double dt, int ev, bool fuse_next);
void reset_model(Ctx& c);  // before loading a new model into a used context
void set_units(Ctx& c, int units);  // CAC_HIP_UNITS_LJ or CAC_HIP_UNITS_METAL constants
void compact_live(Ctx& c);  // dense live-site records from the full ones (after a prune)
void export_lists(Ctx& c, std::vector<int>& row, std::vector<int>& kind, std::vector<int>& owner,
                  std::vector<int>& ucell, std::vector<int>& img);
void scan_exclusive(Ctx& c, const int* in, int* out, int n, int* total);
// Dense live-site arrays written by scan_live for every flagged site k at its
// index i: descriptors, and the record (FP32 from site32f, or FP64 from site).
struct LiveOut {
  const int *sdesc, *simg, *scell;
  const double4* site;
  const float4* site32f;
  int *lsdesc, *lsimg, *lscell;
  float4* site32;
  double4* site_live;
  int mixed;
  unsigned char* clear;  // the flags, zeroed once read (ready for the next prune)
};
// idx = exclusive scan of the 0/1 flags, *d_total (device) = their sum, and the
// flagged sites compacted into lo's dense arrays, in one pass after the sums
void scan_live(Ctx& c, const unsigned char* flag, int* idx, int n, int* d_total, const LiveOut& lo);
// stable sort of the candidate sites by cell: sort_keys / sort_vals, cell_start
void sort_sites_by_cell(Ctx& c, const int* tcell, const unsigned* tval, int S);
void site_positions(Ctx& c);  // FP64 records of every site from its descriptor (list build)
void order_rows(Ctx& c);      // rows in Morton order of their cells

// pair.hip
UcellGrid ucell_grid(const Ctx& c);  // unit-cell grid of the element type, for hex_shape_at
// positions of every live site; full = also the FP64 records (always kept in
// FP64 mode; the mixed path needs them only to build or prune lists)
void site_refresh(Ctx& c, bool full, Gate g = Gate());
void pair_compute(Ctx& c, int eflag, int vflag);      // passes + nodal forces + energy/virial sums
void pair_passes(Ctx& c, int eflag, int vflag, Gate g);  // the pair passes alone
void ensure_inner(Ctx& c);                               // re-prune if precision or delta changed
// the pair passes (no energy / virial) of up to max_segments() models that
// segmentable() accepts and that share one potential, one launch per pass on
// `stream`; the caller orders the models' streams around it
bool segmentable(const Ctx& c);
int max_segments();
void pair_passes_segmented(Ctx* const* cs, int n, hipStream_t stream);
void pair_sums(Ctx& c, int eflag, int vflag);            // energy / virial of the last passes

// integrate.hip
void nve_initial(Ctx& c, double dt);
void nve_final(Ctx& c, double dt);
// speculative step: kick + drift + both displacement checks in one pass, and
// nodal force assembly + final kick in one pass
void step_begin(Ctx& c, double dt, int step, bool outer);
void step_end(Ctx& c, double dt, Gate g);
// step_end of step N fused with step_begin of next_step = N + 1 (no host read
// of the full-step velocities in between); records next_step in begun_step
void step_end_begin(Ctx& c, double dt, int next_step, bool outer);
void update_centers(Ctx& c);
void nodal_assembly(Ctx& c);
void thermo(Ctx& c, double* ke, double* temp);
// thermo() with no device allocation: the sums go through the file-scope
// scratch array of integrate.hip (c.scratch, its address on c's device)
void thermo_noalloc(Ctx& c, double* ke, double* temp);
// that array's device address on `device` (current), taken once per device
double* scratch_on(int device);
void reduce_rows(Ctx& c, const double* data, int n, int ncomp, double* out);

// timing helpers (api.hip)
// GPU time between two events on the context stream; sections may nest.
// kernel_bound: the section is one kernel launched with hipExtLaunchKernel,
// which takes the two events and gives them the kernel's own start and end
// times (a plain event pair also spans the dispatch of the launch and its
// completion signal); the caller launches with start / stop when they are set.
struct ScopedTimer {
  Ctx& c;
  const char* name;
  bool bound = false;
  hipEvent_t start = nullptr, stop = nullptr;
  ScopedTimer(Ctx& cc, const char* nm, bool kernel_bound = false);
  ~ScopedTimer();
};
// waits for the stream and adds every recorded section to its timer
void resolve_timers(Ctx& c);

inline int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

}  // namespace cac

// the C ABI's opaque handle (include/cac_hip.h)
struct cac_hip_ctx {
  cac::Ctx c;
};
