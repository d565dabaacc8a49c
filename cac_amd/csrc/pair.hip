// pair.hip -- site refresh and the EAM (three passes) and LJ force kernels.
//
// Replaces PairEAM::compute (pair_eam.cpp:618-1032), PairEAM::compute_fp
// (pair_eam.cpp:1160-1169), PairLJCut::compute (pair_lj_cut.cpp:227-417) and the
// forward/reverse comm they rely on (CommBrick::forward_comm / reverse_comm,
// forward_comm_pair / reverse_comm_pair).
//
// Every point (atom or Gauss point) owns a row of the full neighbor list and
// sums its own density, force, energy and virial from that row: no scatter, no
// float atomics, a fixed summation tree, so results are bit-reproducible.  The
// reference's half list stores a pair once and applies the reaction to j only
// when j is an atom or a Gauss point; the full list yields the same forces,
// because every site that would receive a reaction is itself a row, and sites
// that would not (non-Gauss interpolated atoms) own no row.  Energies and
// virials carry the reference's per-side scale (1/2 for an atom, nodal_weight/2
// for a node, 0 otherwise; pair_eam.cpp:867-918, 979-984).
//
// Work layout: TPR lanes share a row; lane l takes entries l, l+TPR, ... of the
// row-major list, so one gather instruction of a wavefront touches few cache
// lines (consecutive entries are sites of the same cells), and the lanes'
// partial sums are combined with a fixed xor-shuffle tree.  Workgroups are
// persistent and XCD-aware: workgroup b runs on XCD b % 8 and walks a
// contiguous eighth of the cell-ordered rows, so neighbouring rows share that
// XCD's L2.
//
// Mixed precision: distances come from FP64 positions and are rounded to FP32;
// r, the spline evaluations (FP32 tables staged in LDS) and the pair term run
// in FP32; density, force, energy and virial accumulate in FP64.
#include <hip/hip_ext.h>

#include "cac_ctx.hpp"

namespace cac {

constexpr int PB = 512;  // threads per workgroup of the pair kernels
constexpr int NXCD = 8;

struct PairArgs {
  const double4* site;
  const float4* site32;
  const int* row_site;
  const int* row_point;
  const int* nnum;
  const unsigned* nlist;
  int R, cap, A;
  double L[3];
  float cs[3];
  double cutsq;
  long nchunks;
};

// A mixed-path list entry is site | c << 25: the site image lies d = (dx, dy,
// dz) cells from the row's cell, each component in -2..2, and
// c = (dx + 2) + 5 (dy + 2) + 25 (dz + 2).
constexpr int MIX_SHIFT = 25;
constexpr unsigned MIX_MASK = (1u << MIX_SHIFT) - 1u;

// mixed-path pair vector: x_i - x_j = (u_i - u_j) - d * cell edge with u
// relative to each site's cell corner.  d * edge is exact in FP32 (|d| <= 2),
// so each component rounds once.  d is decoded from c in FP32 with no memory
// access: dz = floor((c + 1/2) / 25) - 2, then t = c - 25 dz = (dx + 2) +
// 5 (dy + 2) + 50, dy = floor((t - 50 + 1/2) / 5) - 2 and dx = t - 5 dy - 62.
// Every quotient lies at least 0.02 from an integer, far beyond the rounding
// of the FP32 multiply-adds, so the floors are exact.
__device__ __forceinline__ void mixed_delta(unsigned ent, const float4& pi, const float4& sj, const float* cs,
                                            float& dx, float& dy, float& dz) {
  const float c = (float)(ent >> MIX_SHIFT);
  const float oz = __builtin_floorf(__builtin_fmaf(c, 0.04f, -1.98f));
  const float t = __builtin_fmaf(oz, -25.0f, c);
  const float oy = __builtin_floorf(__builtin_fmaf(t, 0.2f, -11.9f));
  const float ox = __builtin_fmaf(oy, -5.0f, t) - 62.0f;
  dx = (pi.x - sj.x) - ox * cs[0];
  dy = (pi.y - sj.y) - oy * cs[1];
  dz = (pi.z - sj.z) - oz * cs[2];
}

__device__ __forceinline__ void img_shift(unsigned code, const double* L, double& sx, double& sy, double& sz) {
  sx = (double)((int)(code % 3u) - 1) * L[0];
  sy = (double)((int)((code / 3u) % 3u) - 1) * L[1];
  sz = (double)((int)(code / 9u) - 1) * L[2];
}

// XCD-aware persistent schedule: chunk c of rows is handled by the workgroups
// of XCD x = c / per_xcd, in order.
struct Sched {
  long first, end, stride;
};

__device__ __forceinline__ Sched sched(long nchunks) {
  const int nwg = gridDim.x;
  const int xcd = blockIdx.x % NXCD;
  const int wpx = nwg / NXCD;
  const int idx = blockIdx.x / NXCD;
  const long per = (nchunks + NXCD - 1) / NXCD;
  Sched s;
  s.first = (long)xcd * per + idx;
  s.end = (long)(xcd + 1) * per < nchunks ? (long)(xcd + 1) * per : nchunks;
  s.stride = wpx;
  return s;
}

// Sum over the TPR lanes of a row with DPP lane moves (quad swaps, then the
// half-row and row mirrors): a fixed tree, no LDS traffic, and every lane of
// the row ends with the same total.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int TPR>
__device__ __forceinline__ double row_sum(double v) {
  if (TPR >= 2) v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  if (TPR >= 4) v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  if (TPR >= 8) v += dpp_move<0x141>(v);  // row_half_mirror
  if (TPR >= 16) v += dpp_move<0x140>(v); // row_mirror
  if (TPR >= 32) v += __shfl_xor(v, 16, 32);
  return v;
}


// ------------------------------------------------------------------ site refresh
// forward_comm equivalent: every live site's position from its source point,
// or from the shape-function interpolation of its element's nodes, plus its
// wrap shift.  One thread per site in site order, so the record writes are
// coalesced; the node and shape-function reads of neighbouring sites (same
// cell, mostly the same element) hit the cache.  The mixed path keeps a FP32
// copy relative to the corner of the site's neighbor cell; its FP64 records
// are written only when a list is built or pruned.

struct Frame {
  double L[3], lo[3], cs[3];
};

// W64: 0 = FP32 copy only, 1 = FP64 records too, 2 = FP64 records too when
// the step re-prunes the inner list
// Only the sites the inner lists reference (stamped by the last prune) are
// refreshed, except on a step that builds or prunes lists, which reads all.
// Two sites per thread (k and k + 256 of the block's range), every
// descriptor load issued before any is used, so that the two sites' memory
// round trips overlap.
//
// Full mode (a neighbor build, or a step that re-prunes the inner list)
// refreshes every site's record the prune reads (FP32 cell-relative in the
// mixed path, FP64 otherwise); the dense live records then come from them
// (compact_live).
// Otherwise only the live sites are refreshed, straight into their dense
// records: FP32 cell-relative in the mixed path, FP64 otherwise.
constexpr int RSPT = 2;

template <bool MIXED>
__global__ __launch_bounds__(256) void k_refresh(const double* __restrict__ x, UcellGrid ug, Frame fr, int S, int A,
                                                 int P, const int* __restrict__ sdesc, const int* __restrict__ simg,
                                                 const int* __restrict__ scell, double4* __restrict__ site,
                                                 float4* __restrict__ site32f, const int* __restrict__ nlive,
                                                 const int* __restrict__ lsdesc, const int* __restrict__ lsimg,
                                                 const int* __restrict__ lscell, float4* __restrict__ site32,
                                                 double4* __restrict__ site_live, int all, Gate g) {
#pragma clang fp contract(off)
  if (g.skip()) return;
  const bool full = all || (g.step >= 0 && g.prune());
  const int n = full ? S : *nlive;
  // the grid covers every site; past the live ones a dense refresh has no work
  if (blockIdx.x * blockDim.x * RSPT >= n) return;
  const int* dsc = full ? sdesc : lsdesc;
  const int* img = full ? simg : lsimg;
  const int k0 = blockIdx.x * blockDim.x * RSPT + threadIdx.x;
  int kk[RSPT], d[RSPT], im[RSPT], cc[RSPT];
#pragma unroll
  for (int t = 0; t < RSPT; t++) {
    kk[t] = k0 + t * blockDim.x;
    const int kc = kk[t] < n ? kk[t] : 0;
    d[t] = dsc[kc];
    im[t] = img[kc];
    cc[t] = MIXED ? (full ? scell[kc] : lscell[kc]) : 0;
  }
#pragma unroll
  for (int t = 0; t < RSPT; t++) {
    if (kk[t] >= n) continue;
    double y[3];
    if (d[t] >= 0) {
      for (int c = 0; c < 3; c++) y[c] = x[(size_t)c * P + d[t]];
    } else {
      int e, u;
      split_ucell(ug, -d[t] - 1, e, u);
      double N[NPE];
      hex_shape_at(ug, u, N);
      const size_t base = (size_t)A + (size_t)NPE * e;
      for (int c = 0; c < 3; c++) {
        const double* nx = x + (size_t)c * P + base;
        double s = 0.0;
        for (int q = 0; q < NPE; q++) s += N[q] * nx[q];
        y[c] = s;
      }
    }
    double sh[3];
    img_shift((unsigned)im[t], fr.L, sh[0], sh[1], sh[2]);
    for (int c = 0; c < 3; c++) y[c] = y[c] + sh[c];
    const int k = kk[t];
    if (MIXED) {
      // the full records carry the site's packed cell in w, with the top bit
      // set for a site that is a point (the prune reads both there); the
      // dense ones get F'(rho) in w from the embedding pass
      const unsigned wfull = (unsigned)cc[t] | (d[t] >= 0 ? POINT_SITE_BIT : 0u);
      const float4 r32 = make_float4((float)(y[0] - (fr.lo[0] + cell_ix(cc[t]) * fr.cs[0])),
                                     (float)(y[1] - (fr.lo[1] + cell_iy(cc[t]) * fr.cs[1])),
                                     (float)(y[2] - (fr.lo[2] + cell_iz(cc[t]) * fr.cs[2])),
                                     full ? __uint_as_float(wfull) : 0.0f);
      (full ? site32f : site32)[k] = r32;
    } else {
      (full ? site : site_live)[k] = make_double4(y[0], y[1], y[2], 0.0);
    }
  }
}

UcellGrid ucell_grid(const Ctx& c) {
  UcellGrid ug;
  for (int d = 0; d < 3; d++) ug.h[d] = c.ncell[d] > 1 ? 2.0 / (c.ncell[d] - 1) : 0.0;
  ug.ny = c.ncell[1] > 0 ? c.ncell[1] : 1;
  ug.nz = c.ncell[2] > 0 ? c.ncell[2] : 1;
  ug.nucell = c.nucell > 0 ? c.nucell : 1;
  ug.inv_nucell = 1.0 / ug.nucell;
  ug.inv_nyz = 1.0f / (float)(ug.ny * ug.nz);
  ug.inv_nz = 1.0f / (float)ug.nz;
  return ug;
}

static Frame frame_of(const Ctx& c) {
  Frame fr;
  for (int d = 0; d < 3; d++) {
    fr.L[d] = c.L[d];
    fr.lo[d] = c.lo[d];
    fr.cs[d] = c.L[d] / c.nc[d];
  }
  return fr;
}

// dense records: every live site, plus the dummy record the padded list rows name
static void reserve_live(Ctx& c) {
  c.lsdesc.reserve(c.S);
  c.lsimg.reserve(c.S);
  c.lscell.reserve(c.S);
  if (c.precision == 1) {
    c.site32.reserve((size_t)c.S + 1);
    c.site32f.reserve(c.S);
  } else {
    c.site_live.reserve((size_t)c.S + 1);
  }
}

void site_refresh(Ctx& c, bool full, Gate g) {
  ScopedTimer tm(c, "site_refresh");
  c.x_dirty = false;
  const bool mixed = c.precision == 1;
  reserve_live(c);
  c.d_nlive.reserve(1);
  const Frame fr = frame_of(c);
  const UcellGrid ug = ucell_grid(c);
  const int grid = grid_for(c.S, 256 * RSPT);
  auto kr = mixed ? k_refresh<true> : k_refresh<false>;
  kr<<<grid, 256, 0, c.stream>>>(c.x.get(), ug, fr, c.S, c.A, c.P, c.sdesc.get(), c.simg.get(), c.scell.get(),
                                 c.site.get(), c.site32f.get(), c.d_nlive.get(), c.lsdesc.get(), c.lsimg.get(),
                                 c.lscell.get(), c.site32.get(), c.site_live.get(), full ? 1 : 0, g);
}

void site_positions(Ctx& c) {
  const Frame fr = frame_of(c);
  const UcellGrid ug = ucell_grid(c);
  k_refresh<false><<<grid_for(c.S, 256 * RSPT), 256, 0, c.stream>>>(
      c.x.get(), ug, fr, c.S, c.A, c.P, c.sdesc.get(), c.simg.get(), c.scell.get(), c.site.get(), nullptr, nullptr,
      nullptr, nullptr, nullptr, nullptr, nullptr, 1, Gate());
}

// dense live records from the full ones, after a prune flagged the live sites:
// live_idx, the live count (device) and the dense arrays in one scan
void compact_live(Ctx& c) {
  reserve_live(c);
  c.live_idx.reserve(c.S);
  c.d_nlive.reserve(1);
  LiveOut lo{c.sdesc.get(), c.simg.get(), c.scell.get(), c.site.get(), c.site32f.get(), c.lsdesc.get(),
             c.lsimg.get(), c.lscell.get(), c.site32.get(), c.site_live.get(), c.precision == 1 ? 1 : 0,
             c.live_flag.get()};
  scan_live(c, c.live_flag.get(), c.live_idx.get(), c.S, c.d_nlive.get(), lo);
  c.live_flag_clean = std::max(c.live_flag_clean, (size_t)c.S);
}

// ------------------------------------------------------------------ spline helpers
// knot index and fraction as pair_eam.cpp:799-805 and 1160-1169

template <typename T>
__device__ __forceinline__ int knot_r(T r, T rdr, int nr, T& p) {
  p = r * rdr + (T)1.0;
  int m = (int)p;
  m = m < nr - 1 ? m : nr - 1;
  m = m > 1 ? m : 1;  // r >= 0 gives m >= 1; this only keeps a non-finite r inside the table
  p -= (T)m;
  p = p < (T)1.0 ? p : (T)1.0;
  return m;
}

// FP32 knot of the mixed path: m = int(r rdr + 1), p its fraction, as the
// reference forms them for r < (nr - 1) dr.  Past that (the last knot
// interval, where the tabulated functions of a potential that reaches zero
// at its cutoff are near zero) the reference clamps to knot nr - 1 at p = 1;
// here knot nr continues it (the density table linearly, the force tables
// with the clamped value), and m is kept at most nr + 1 (the all-zero knot),
// so a table read stays inside the table.  An entry beyond the cutoff takes
// the zero knot: the caller's `inside`.
__device__ __forceinline__ int knot_r32(float r, float rdr, int nr, bool inside, float& p) {
  const float x = r * rdr + 1.0f;
  p = __builtin_amdgcn_fractf(x);
  int m = (int)x;
  m = m < nr + 1 ? m : nr + 1;
  return inside ? m : nr + 1;
}

__device__ __forceinline__ int knot_rho(double rho, double rdrho, int nrho, double& p) {
  p = rho * rdrho + 1.0;
  int m = (int)p;
  m = m < nrho - 1 ? m : nrho - 1;
  m = m > 1 ? m : 1;
  p -= m;
  p = p < 1.0 ? p : 1.0;
  return m;
}

// spline tables: knot m holds two vectors at 2m and 2m + 1 in device memory
// (FP64: rhor c3..c6 and z2r c3..c6; mixed: z2r c3..c6, then rhor's
// derivative coefficients c0..c2 and z2r's c2); derivatives are formed from the
// value coefficients, ((3 c3 p + 2 c4) p + c5) / dr, which is what the
// reference's coefficients c0..c2 hold.  The density pass stages the rhor
// half alone; the force pass stages both halves as two tables of n vectors
// each, so the lanes' reads of unrelated knots spread over all LDS banks (a
// 16-byte stride for the mixed path) rather than every other bank group.
template <typename V>
__device__ __forceinline__ void stage_table(V* lds, const V* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
}

template <typename V>
__device__ __forceinline__ void stage_halves(V* lds, const V* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < 2 * n; i += blockDim.x) lds[(i & 1) * n + (i >> 1)] = g[i];
  __syncthreads();
}

template <typename V>
__device__ __forceinline__ void stage_table_strided(V* lds, const V* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[2 * i];
  __syncthreads();
}

// ------------------------------------------------------------------ row walker
// The list build pads every inner row to a multiple of ROW_PAD (16) entries
// with an entry naming a dummy site far from every point, and the passes walk
// the whole padded row: no index clamp and no row-end mask.  An entry beyond
// the cutoff (a dummy among them) evaluates the all-zero spline knot nr + 1
// appended to the tables, so it contributes an exact zero.
//
// Lane l of a row's TPR lanes takes entries l, l + TPR, ...: full batches of
// U entries per lane (U list loads, then U site gathers, then U pair terms,
// so U independent memory chains are in flight), then what is left of the
// padded row in one batch each of U/2, U/4, ... 2 entries per lane, as far
// as they fit.  Short batches leave few memory chains in flight, so a row is
// walked to its padding (a batch of dummies costs less than a short tail:
// rounding rows to 8 entries made the force pass 4 % slower, a last batch
// of 1 entry per lane slower still), and the remainder takes the largest
// batches that fit.  `flush` runs after every GROUP entries of a
// lane and after its last one (the passes move their FP32 group sums into
// FP64 sums there); groups follow the lane's entry count, not the batches, so
// the sums do not depend on U.
constexpr int UNROLL = 4;  // entries per lane in flight in a full batch
constexpr int GROUP = 4;   // entries of a lane summed in FP32 before the FP64 sum

// One level of a row walk: batches of UB entries per lane while a whole batch
// fits in the padded row, then the level with half as many, down to 2.
template <int TPR, int UB, typename Site, typename Fetch, typename Body, typename Flush>
__device__ __forceinline__ void walk_level(const unsigned* __restrict__ rp, int& j, int np, int& cnt, Fetch fetch,
                                           Body body, Flush flush) {
  static_assert(ROW_PAD % (2 * TPR) == 0, "the last level must stay inside the padded row");
  for (; j + UB * TPR <= np; j += UB * TPR) {
    unsigned w[UB];
#pragma unroll
    for (int u = 0; u < UB; u++) w[u] = rp[j + u * TPR];
    Site sj[UB];
#pragma unroll
    for (int u = 0; u < UB; u++) sj[u] = fetch(w[u]);
#pragma unroll
    for (int u = 0; u < UB; u++) {
      body(w[u], sj[u]);
      // the levels above this one leave cnt a multiple of GROUP
      if (UB % GROUP == 0 ? (u + 1) % GROUP == 0 : (cnt + u + 1) % GROUP == 0) flush();
    }
    cnt += UB;
  }
  if constexpr (UB > 2) walk_level<TPR, UB / 2, Site>(rp, j, np, cnt, fetch, body, flush);
}

template <int TPR, int U, typename Site, typename Fetch, typename Body, typename Flush>
__device__ __forceinline__ void walk_row(const unsigned* __restrict__ base, unsigned row0, int n, int lane,
                                         Fetch fetch, Body body, Flush flush) {
  static_assert(U % GROUP == 0 && (U & (U - 1)) == 0, "a full batch holds whole groups; levels halve it");
  const unsigned* __restrict__ rp = base + row0 + lane;
  const int np = (n + ROW_PAD - 1) / ROW_PAD * ROW_PAD;
  int j = 0, cnt = 0;
  walk_level<TPR, U, Site>(rp, j, np, cnt, fetch, body, flush);
  if (cnt % GROUP) flush();
}

// ------------------------------------------------------------------ EAM density (pair_eam.cpp:721-813)

// Density of the rows of chunk ch (NT / TPR rows, TPR lanes each) of one
// model; the spline table is staged in LDS by the caller.
template <int TPR, bool MIXED, int NT, int U>
__device__ __forceinline__ void density_chunk(const PairArgs& a, long ch, const float4* tab32,
                                              const double4* lds64, double rdr, int nr,
                                              double* __restrict__ rho_out) {
  const int lane = threadIdx.x % TPR;
  const int rsub = threadIdx.x / TPR;
  const float rdr32 = (float)rdr;
  const float cut32 = (float)a.cutsq;
  const int r = (int)(ch * (NT / TPR)) + rsub;
  double acc = 0.0;
  if (r < a.R) {
    const int k = a.row_site[r];
    const int n = a.nnum[r];
    const unsigned row0 = (unsigned)r * (unsigned)a.cap;
    if (MIXED) {
      const float4 pi = a.site32[k];
      float acc32 = 0.0f;
      walk_row<TPR, U, float4>(
          a.nlist, row0, n, lane, [&](unsigned e) { return a.site32[e & MIX_MASK]; },
          [&](unsigned ent, const float4& sj) {
            float dx, dy, dz;
            mixed_delta(ent, pi, sj, a.cs, dx, dy, dz);
            const float rsq = dx * dx + dy * dy + dz * dz;
            float p;
            // beyond rc (or the dummy): the zero knot
            const int m = knot_r32(__builtin_amdgcn_sqrtf(rsq), rdr32, nr, rsq < cut32, p);
            const float4 c = tab32[m];
            acc32 += ((c.x * p + c.y) * p + c.z) * p + c.w;
          },
          [&]() {
            acc += (double)acc32;  // a lane's group of at most four terms, summed in FP32
            acc32 = 0.0f;
          });
    } else {
      const double4 pi = a.site[k];
      walk_row<TPR, U, double4>(
          a.nlist, row0, n, lane, [&](unsigned e) { return a.site[e & SITE_MASK]; },
          [&](unsigned ent, const double4& sj) {
            double shx, shy, shz;
            img_shift(ent >> IMG_SHIFT, a.L, shx, shy, shz);
            const double delx = pi.x - (sj.x + shx);
            const double dely = pi.y - (sj.y + shy);
            const double delz = pi.z - (sj.z + shz);
            const double rsq = delx * delx + dely * dely + delz * delz;
            double p;
            int m = knot_r<double>(sqrt(rsq), rdr, nr, p);
            m = rsq < a.cutsq ? m : nr + 1;
            const double4 c = lds64[m];
            acc += ((c.x * p + c.y) * p + c.z) * p + c.w;
          },
          []() {});
    }
  }
  acc = row_sum<TPR>(acc);
  if (lane == 0 && r < a.R) rho_out[a.row_point[r]] = acc;
}

// the density pass stages only the rho(r) half of the table (knot m at m)
template <int TPR, bool MIXED, int NT, int U>
__global__ __launch_bounds__(NT) void k_eam_density(PairArgs a, const void* __restrict__ tab, double rdr, int nr,
                                                    double* __restrict__ rho_out, Gate g) {
  if (g.skip()) return;
  extern __shared__ float4 lds_raw[];
  float4* tab32 = lds_raw;
  double4* lds64 = reinterpret_cast<double4*>(lds_raw);
  if (MIXED) {
    stage_table(tab32, static_cast<const float4*>(tab), nr + 2);
  } else {
    stage_table_strided(lds64, static_cast<const double4*>(tab), nr + 2);
  }
  const Sched s = sched(a.nchunks);
  for (long ch = s.first; ch < s.end; ch += s.stride)
    density_chunk<TPR, MIXED, NT, U>(a, ch, tab32, lds64, rdr, nr, rho_out);
}

// ------------------------------------------------------------------ embedding derivative per site
// atoms and nodes: F'(rho) of their own density (pair_eam.cpp:822-832, 910);
// interpolated atoms: F'(sum_k N_k rho_node_k), the value the reference
// recomputes for every pair (pair_eam.cpp:959-969)

__device__ __forceinline__ double fprime(const double* __restrict__ frho, double rho, double rdrho, int nrho) {
  double p;
  const int m = knot_rho(rho, rdrho, nrho, p);
  const double* c = frho + (size_t)m * 7;
  return (c[0] * p + c[1]) * p + c[2];
}

// F'(rho) of every live site (compute_fp, pair_eam.cpp:1160-1169): a point's
// own density, or the shape-function interpolation of its element's nodal
// densities for an interpolated atom (pair_eam.cpp:700-760); two live sites
// per thread in dense order.  Only the F' word of each record is written.
template <bool MIXED>
__device__ __forceinline__ void embed_block(int blk, const double* __restrict__ rho, const UcellGrid& ug,
                                            const int* __restrict__ lsdesc, const int* __restrict__ nlive, int A,
                                            const double* __restrict__ frho, double rdrho, int nrho,
                                            double4* __restrict__ site_live, float4* __restrict__ site32) {
#pragma clang fp contract(off)
  const int n = *nlive;
  if (blk * blockDim.x * RSPT >= n) return;  // the grid covers every site, not only the live ones
  const int k0 = blk * blockDim.x * RSPT + threadIdx.x;
  int kk[RSPT], d[RSPT];
#pragma unroll
  for (int t = 0; t < RSPT; t++) {
    kk[t] = k0 + t * blockDim.x;
    d[t] = lsdesc[kk[t] < n ? kk[t] : 0];
  }
#pragma unroll
  for (int t = 0; t < RSPT; t++) {
    if (kk[t] >= n) continue;
    double rj;
    if (d[t] >= 0) {
      rj = rho[d[t]];
    } else {
      int e, u;
      split_ucell(ug, -d[t] - 1, e, u);
      double N[NPE];
      hex_shape_at(ug, u, N);
      const double* nr = rho + (size_t)A + (size_t)NPE * e;
      rj = 0.0;
      for (int q = 0; q < NPE; q++) rj += N[q] * nr[q];
    }
    const double fp = fprime(frho, rj, rdrho, nrho);
    if (MIXED)
      reinterpret_cast<float*>(site32 + kk[t])[3] = (float)fp;
    else
      reinterpret_cast<double*>(site_live + kk[t])[3] = fp;
  }
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_embed(const double* __restrict__ rho, UcellGrid ug,
                                               const int* __restrict__ lsdesc, const int* __restrict__ nlive, int A,
                                               const double* __restrict__ frho, double rdrho, int nrho,
                                               double4* __restrict__ site_live, float4* __restrict__ site32) {
  embed_block<MIXED>(blockIdx.x, rho, ug, lsdesc, nlive, A, frho, rdrho, nrho, site_live, site32);
}

// ------------------------------------------------------------------ EAM force (pair_eam.cpp:853-1027)

struct EamForceArgs {
  const void* tab;     // LDS-staged spline coefficients, two float4 or double4 per knot
  const double* frho;  // [(nrho+1)*7]
  const double* rho;   // per point
  double rdr, rdrho, rhomax;
  int nr, nrho;
  double escale_node, vscale_node;
  // per-point energy and virial [7][P] (Pair::eatom / enode, vatom / vnode),
  // or null: each point's F(rho) plus half of its pair energies, and half of
  // its pair virials (pair.cpp:441-476, pair_eam.cpp:863-918, 979-984)
  double* peratom;
};

// Force (and, with EV, energy and virial) of the rows of chunk ch of one
// model; as density_chunk.
template <int TPR, bool MIXED, bool EV, int NT, int U>
__device__ __forceinline__ void force_chunk(const PairArgs& a, const EamForceArgs& q, long ch, const float4* tab32, const double4* lds64, double* __restrict__ f,
                                            double* __restrict__ gf, int P, int E8, double* __restrict__ erow) {
  const int lane = threadIdx.x % TPR;
  const int rsub = threadIdx.x / TPR;
  const float rdr32 = (float)q.rdr;
  const int r = (int)(ch * (NT / TPR)) + rsub;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  double ephi = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
  int k = 0;
  if (r < a.R) {
    k = a.row_site[r];
    const int n = a.nnum[r];
    const unsigned row0 = (unsigned)r * (unsigned)a.cap;
    if (MIXED) {
      const float4 pi = a.site32[k];
      const float cut32 = (float)a.cutsq;
      float fx32 = 0.0f, fy32 = 0.0f, fz32 = 0.0f;
      walk_row<TPR, U, float4>(
          a.nlist, row0, n, lane, [&](unsigned e) { return a.site32[e & MIX_MASK]; },
          [&](unsigned ent, const float4& sj) {
            float dx, dy, dz;
            mixed_delta(ent, pi, sj, a.cs, dx, dy, dz);
            const float rsq = dx * dx + dy * dy + dz * dz;
            float p;
            // beyond rc (or the dummy): the zero knot
            const int m = knot_r32(__builtin_amdgcn_sqrtf(rsq), rdr32, q.nr, rsq < cut32, p);
            const float4 ca = tab32[m];             // A = -rho'(r) / r
            const float4 cb = tab32[q.nr + 2 + m];  // B = -phi'(r) / r
            const float fa = ((ca.x * p + ca.y) * p + ca.z) * p + ca.w;
            const float fb = ((cb.x * p + cb.y) * p + cb.z) * p + cb.w;
            const float fpair = __builtin_fmaf(pi.w + sj.w, fa, fb);
            fx32 = __builtin_fmaf(dx, fpair, fx32);
            fy32 = __builtin_fmaf(dy, fpair, fy32);
            fz32 = __builtin_fmaf(dz, fpair, fz32);
            if (EV) {
              const float4 cc = tab32[2 * (q.nr + 2) + m];  // C = phi(r)
              const float phi = ((cc.x * p + cc.y) * p + cc.z) * p + cc.w;
              ephi += (double)phi;
              v0 += (double)(dx * dx * fpair);
              v1 += (double)(dy * dy * fpair);
              v2 += (double)(dz * dz * fpair);
              v3 += (double)(dx * dy * fpair);
              v4 += (double)(dx * dz * fpair);
              v5 += (double)(dy * dz * fpair);
            }
          },
          [&]() {
            // a lane's group of at most four pair forces is summed in FP32,
            // the point's force in FP64
            fx += (double)fx32;
            fy += (double)fy32;
            fz += (double)fz32;
            fx32 = fy32 = fz32 = 0.0f;
          });
    } else {
      const double4 pi = a.site[k];
      walk_row<TPR, U, double4>(
          a.nlist, row0, n, lane, [&](unsigned e) { return a.site[e & SITE_MASK]; },
          [&](unsigned ent, const double4& sj) {
            double shx, shy, shz;
            img_shift(ent >> IMG_SHIFT, a.L, shx, shy, shz);
            const double delx = pi.x - (sj.x + shx);
            const double dely = pi.y - (sj.y + shy);
            const double delz = pi.z - (sj.z + shz);
            const double rsq = delx * delx + dely * dely + delz * delz;
            const double rr = sqrt(rsq);
            double p;
            int m = knot_r<double>(rr, q.rdr, q.nr, p);
            m = rsq < a.cutsq ? m : q.nr + 1;  // beyond rc (or the dummy): the zero knot
            const double4 cr = lds64[m];
            const double4 cz = lds64[q.nr + 2 + m];
            const double rhop = ((3.0 * cr.x * p + 2.0 * cr.y) * p + cr.z) * q.rdr;
            const double z2p = ((3.0 * cz.x * p + 2.0 * cz.y) * p + cz.z) * q.rdr;
            const double z2 = ((cz.x * p + cz.y) * p + cz.z) * p + cz.w;
            const double recip = 1.0 / rr;
            const double phi = z2 * recip;
            const double phip = z2p * recip - phi * recip;
            const double psip = pi.w * rhop + sj.w * rhop + phip;
            const double fpair = -1.0 * psip * recip;
            fx += delx * fpair;
            fy += dely * fpair;
            fz += delz * fpair;
            if (EV) {
              ephi += phi;
              v0 += delx * delx * fpair;
              v1 += dely * dely * fpair;
              v2 += delz * delz * fpair;
              v3 += delx * dely * fpair;
              v4 += delx * delz * fpair;
              v5 += dely * delz * fpair;
            }
          },
          []() {});
    }
  }
  fx = row_sum<TPR>(fx);
  fy = row_sum<TPR>(fy);
  fz = row_sum<TPR>(fz);
  if (EV) {
    ephi = row_sum<TPR>(ephi);
    v0 = row_sum<TPR>(v0);
    v1 = row_sum<TPR>(v1);
    v2 = row_sum<TPR>(v2);
    v3 = row_sum<TPR>(v3);
    v4 = row_sum<TPR>(v4);
    v5 = row_sum<TPR>(v5);
  }
  if (lane == 0 && r < a.R) {
    const int pnt = a.row_point[r];
    const bool is_atom = pnt < a.A;
    if (is_atom) {
      f[pnt] = fx;
      f[(size_t)P + pnt] = fy;
      f[2 * (size_t)P + pnt] = fz;
    } else {
      const int g = pnt - a.A;
      gf[g] = fx;
      gf[(size_t)E8 + g] = fy;
      gf[2 * (size_t)E8 + g] = fz;
    }
    if (EV) {
      // embedding energy F(rho) at atoms (weight 1) and nodes (weight nodal_weight)
      const double rho = q.rho[pnt];
      const double fpi = fprime(q.frho, rho, q.rdrho, q.nrho);
      double p;
      const int m = knot_rho(rho, q.rdrho, q.nrho, p);
      const double* c = q.frho + (size_t)m * 7;
      double phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
      if (rho > q.rhomax) phi += fpi * (rho - q.rhomax);
      const double es = is_atom ? 0.5 : q.escale_node;
      const double vs = is_atom ? 0.5 : q.vscale_node;
      erow[r] = phi * (is_atom ? 1.0 : 2.0 * q.escale_node) + es * ephi;
      erow[(size_t)a.R * 1 + r] = vs * v0;
      erow[(size_t)a.R * 2 + r] = vs * v1;
      erow[(size_t)a.R * 3 + r] = vs * v2;
      erow[(size_t)a.R * 4 + r] = vs * v3;
      erow[(size_t)a.R * 5 + r] = vs * v4;
      erow[(size_t)a.R * 6 + r] = vs * v5;
      if (q.peratom) {
        double* pa = q.peratom + pnt;
        pa[0] = phi + 0.5 * ephi;
        pa[(size_t)P] = 0.5 * v0;
        pa[2 * (size_t)P] = 0.5 * v1;
        pa[3 * (size_t)P] = 0.5 * v2;
        pa[4 * (size_t)P] = 0.5 * v3;
        pa[5 * (size_t)P] = 0.5 * v4;
        pa[6 * (size_t)P] = 0.5 * v5;
      }
    }
  }
}

template <int TPR, bool MIXED, bool EV, int NT, int U>
__global__ __launch_bounds__(NT) void k_eam_force(PairArgs a, EamForceArgs q, double* __restrict__ f,
                                                  double* __restrict__ gf, int P, int E8,
                                                  double* __restrict__ erow, Gate g) {
  if (g.skip()) return;
  extern __shared__ float4 lds_raw[];
  float4* tab32 = lds_raw;
  double4* lds64 = reinterpret_cast<double4*>(lds_raw);
  if (MIXED) {
    stage_table(tab32, static_cast<const float4*>(q.tab), (EV ? 3 : 2) * (q.nr + 2));
  } else {
    stage_halves(lds64, static_cast<const double4*>(q.tab), q.nr + 2);
  }
  const Sched s = sched(a.nchunks);
  for (long ch = s.first; ch < s.end; ch += s.stride)
    force_chunk<TPR, MIXED, EV, NT, U>(a, q, ch, tab32, lds64, f, gf, P, E8, erow);
}

// ------------------------------------------------------------------ several models in one launch
// The density, embedding and force passes of up to MAXSEG mixed-precision EAM
// models that share the potential and the launch shape, one launch each (the
// serving loop's replicas).  The models' chunks of rows (blocks of live sites
// for the embedding) are laid end to end; a workgroup finds the model of each
// chunk from the running counts.  A row is walked by the same lanes in the same
// order as in the model's own launch, so its sums are bitwise those of a lone
// run.

constexpr int MAXSEG = 16;

struct SegRows {
  PairArgs a[MAXSEG];
  double* out[MAXSEG];  // density: rho; force: f
  double* gf[MAXSEG];
  int P[MAXSEG], E8[MAXSEG];
  long first[MAXSEG + 1];  // first chunk of each model; first[n] = all chunks
  int n;
};

struct SegSites {
  const double* rho[MAXSEG];
  UcellGrid ug[MAXSEG];
  const int* lsdesc[MAXSEG];
  const int* nlive[MAXSEG];
  float4* site32[MAXSEG];
  int A[MAXSEG];
  int first[MAXSEG + 1];  // first block of each model
  int n;
};

// the model of chunk ch (the same for every lane of the workgroup)
__device__ __forceinline__ int seg_of(const long* first, int n, long ch) {
  int k = 0;
  while (k + 1 < n && ch >= first[k + 1]) k++;
  return k;
}

template <int TPR, int NT, int U>
__global__ __launch_bounds__(NT) void k_eam_density_seg(const SegRows sg, const float4* __restrict__ tab, double rdr,
                                                        int nr) {
  extern __shared__ float4 lds_raw[];
  float4* tab32 = lds_raw;
  stage_table(tab32, tab, nr + 2);
  const Sched s = sched(sg.first[sg.n]);
  for (long ch = s.first; ch < s.end; ch += s.stride) {
    const int k = seg_of(sg.first, sg.n, ch);
    density_chunk<TPR, true, NT, U>(sg.a[k], ch - sg.first[k], tab32, nullptr, rdr, nr, sg.out[k]);
  }
}

template <int TPR, int NT, int U>
__global__ __launch_bounds__(NT) void k_eam_force_seg(const SegRows sg, EamForceArgs q) {
  extern __shared__ float4 lds_raw[];
  float4* tab32 = lds_raw;
  stage_table(tab32, static_cast<const float4*>(q.tab), 2 * (q.nr + 2));
  const Sched s = sched(sg.first[sg.n]);
  for (long ch = s.first; ch < s.end; ch += s.stride) {
    const int k = seg_of(sg.first, sg.n, ch);
    force_chunk<TPR, true, false, NT, U>(sg.a[k], q, ch - sg.first[k], tab32, nullptr, sg.out[k], sg.gf[k],
                                         sg.P[k], sg.E8[k], nullptr);
  }
}

__global__ __launch_bounds__(256) void k_embed_seg(const SegSites sg, const double* __restrict__ frho, double rdrho,
                                                   int nrho) {
  int k = 0;
  while (k + 1 < sg.n && (int)blockIdx.x >= sg.first[k + 1]) k++;
  embed_block<true>((int)blockIdx.x - sg.first[k], sg.rho[k], sg.ug[k], sg.lsdesc[k], sg.nlive[k], sg.A[k], frho,
                    rdrho, nrho, nullptr, sg.site32[k]);
}

// ------------------------------------------------------------------ LJ (pair_lj_cut.cpp:278-414)

struct LjArgs {
  double lj1, lj2, lj3, lj4, escale_node, vscale_node;
  double* peratom;  // per-point energy and virial [7][P], or null (as EamForceArgs)
};

template <int TPR, bool MIXED, bool EV>
__global__ __launch_bounds__(PB) void k_lj_force(PairArgs a, LjArgs q, double* __restrict__ f,
                                                 double* __restrict__ gf, int P, int E8, double* __restrict__ erow,
                                                 Gate g) {
  if (g.skip()) return;
  const int lane = threadIdx.x % TPR;
  const int rsub = threadIdx.x / TPR;
  const Sched s = sched(a.nchunks);
  const float l1 = (float)q.lj1, l2 = (float)q.lj2, l3 = (float)q.lj3, l4 = (float)q.lj4;
  for (long ch = s.first; ch < s.end; ch += s.stride) {
    const int r = (int)(ch * (PB / TPR)) + rsub;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double ev = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
    if (r < a.R) {
      const int k = a.row_site[r];
      const int n = a.nnum[r];
      const unsigned row0 = (unsigned)r * (unsigned)a.cap;
      if (MIXED) {
        const float4 pi = a.site32[k];
        const float cut32 = (float)a.cutsq;
        float fx32 = 0.0f, fy32 = 0.0f, fz32 = 0.0f;
        walk_row<TPR, UNROLL, float4>(
            a.nlist, row0, n, lane, [&](unsigned e) { return a.site32[e & MIX_MASK]; },
            [&](unsigned ent, const float4& sj) {
              float dx, dy, dz;
              mixed_delta(ent, pi, sj, a.cs, dx, dy, dz);
              const float rsq = dx * dx + dy * dy + dz * dz;
              const float okf = rsq < cut32 ? 1.0f : 0.0f;  // beyond rc (or the dummy): nothing
              const float r2inv = __builtin_amdgcn_rcpf(rsq);
              const float r6inv = r2inv * r2inv * r2inv;
              const float forcelj = r6inv * (l1 * r6inv - l2);
              const float fpair = okf * (forcelj * r2inv);
              fx32 = __builtin_fmaf(dx, fpair, fx32);
              fy32 = __builtin_fmaf(dy, fpair, fy32);
              fz32 = __builtin_fmaf(dz, fpair, fz32);
              if (EV) {
                ev += (double)(okf * (r6inv * (l3 * r6inv - l4)));
                v0 += (double)(dx * dx * fpair);
                v1 += (double)(dy * dy * fpair);
                v2 += (double)(dz * dz * fpair);
                v3 += (double)(dx * dy * fpair);
                v4 += (double)(dx * dz * fpair);
                v5 += (double)(dy * dz * fpair);
              }
            },
            [&]() {
              fx += (double)fx32;
              fy += (double)fy32;
              fz += (double)fz32;
              fx32 = fy32 = fz32 = 0.0f;
            });
      } else {
        const double4 pi = a.site[k];
        walk_row<TPR, UNROLL, double4>(
            a.nlist, row0, n, lane, [&](unsigned e) { return a.site[e & SITE_MASK]; },
            [&](unsigned ent, const double4& sj) {
              double shx, shy, shz;
              img_shift(ent >> IMG_SHIFT, a.L, shx, shy, shz);
              const double delx = pi.x - (sj.x + shx);
              const double dely = pi.y - (sj.y + shy);
              const double delz = pi.z - (sj.z + shz);
              const double rsq = delx * delx + dely * dely + delz * delz;
              const double okf = rsq < a.cutsq ? 1.0 : 0.0;
              const double r2inv = 1.0 / rsq;
              const double r6inv = r2inv * r2inv * r2inv;
              const double forcelj = r6inv * (q.lj1 * r6inv - q.lj2);
              const double fpair = okf * (forcelj * r2inv);
              fx += delx * fpair;
              fy += dely * fpair;
              fz += delz * fpair;
              if (EV) {
                ev += okf * (r6inv * (q.lj3 * r6inv - q.lj4));
                v0 += delx * delx * fpair;
                v1 += dely * dely * fpair;
                v2 += delz * delz * fpair;
                v3 += delx * dely * fpair;
                v4 += delx * delz * fpair;
                v5 += dely * delz * fpair;
              }
            },
            []() {});
      }
    }
    fx = row_sum<TPR>(fx);
    fy = row_sum<TPR>(fy);
    fz = row_sum<TPR>(fz);
    if (EV) {
      ev = row_sum<TPR>(ev);
      v0 = row_sum<TPR>(v0);
      v1 = row_sum<TPR>(v1);
      v2 = row_sum<TPR>(v2);
      v3 = row_sum<TPR>(v3);
      v4 = row_sum<TPR>(v4);
      v5 = row_sum<TPR>(v5);
    }
    if (lane == 0 && r < a.R) {
      const int pnt = a.row_point[r];
      const bool is_atom = pnt < a.A;
      if (is_atom) {
        f[pnt] = fx;
        f[(size_t)P + pnt] = fy;
        f[2 * (size_t)P + pnt] = fz;
      } else {
        const int g = pnt - a.A;
        gf[g] = fx;
        gf[(size_t)E8 + g] = fy;
        gf[2 * (size_t)E8 + g] = fz;
      }
      if (EV) {
        const double es = is_atom ? 0.5 : q.escale_node;
        const double vs = is_atom ? 0.5 : q.vscale_node;
        erow[r] = es * ev;
        erow[(size_t)a.R * 1 + r] = vs * v0;
        erow[(size_t)a.R * 2 + r] = vs * v1;
        erow[(size_t)a.R * 3 + r] = vs * v2;
        erow[(size_t)a.R * 4 + r] = vs * v3;
        erow[(size_t)a.R * 5 + r] = vs * v4;
        erow[(size_t)a.R * 6 + r] = vs * v5;
        if (q.peratom) {
          double* pa = q.peratom + pnt;
          pa[0] = 0.5 * ev;
          pa[(size_t)P] = 0.5 * v0;
          pa[2 * (size_t)P] = 0.5 * v1;
          pa[3 * (size_t)P] = 0.5 * v2;
          pa[4 * (size_t)P] = 0.5 * v3;
          pa[5 * (size_t)P] = 0.5 * v4;
          pa[6 * (size_t)P] = 0.5 * v5;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ orchestration

// ENERGY_ATOM, VIRIAL_ATOM or VIRIAL_CENTROID (force.h:11-19; the centroid
// virial of these pair styles is the two-body one, pair.cpp:280-281)
static bool peratom_wanted(int eflag, int vflag) { return (eflag & 2) || (vflag & 12); }

static PairArgs pair_args(Ctx& c, double cutsq, int rows_per_wg) {
  PairArgs a;
  a.site = c.site_live.get();
  a.site32 = c.site32.get();
  a.row_site = c.row_live.get();
  a.row_point = c.row_point.get();
  a.nnum = c.innum.get();
  a.nlist = reinterpret_cast<const unsigned*>(c.ilist.get());
  a.R = c.P;
  a.cap = c.icap;
  a.A = c.A;
  for (int d = 0; d < 3; d++) {
    a.L[d] = c.L[d];
    a.cs[d] = (float)(c.L[d] / c.nc[d]);
  }
  a.cutsq = cutsq;
  a.nchunks = (c.P + rows_per_wg - 1) / rows_per_wg;
  return a;
}

static int pair_grid(long nchunks, long cap) {
  long g = cap < nchunks ? cap : nchunks;
  g = ((g + NXCD - 1) / NXCD) * NXCD;
  return (int)(g < NXCD ? NXCD : g);
}

// kernels staging more than 64 KB of LDS must say so before their first launch
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) CAC_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// Workgroups of a persistent kernel: as many as can be resident at once
// (registers, LDS and waves per CU as the runtime reports them), so that no
// workgroup waits for another to finish before it starts its share.
template <typename K>
static int resident_grid(Ctx& c, K kernel, int nt, size_t lds, long nchunks) {
  const void* kp = reinterpret_cast<const void*>(kernel);
  int& per_cu = c.resident[{kp, lds}];
  if (per_cu == 0) {
    allow_lds(kernel, lds);
    CAC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, nt, lds));
    if (per_cu < 1) per_cu = 1;
  }
  return pair_grid(nchunks, std::min<long>((long)per_cu * c.ncu, c.grid_cap));
}

template <int TPR, bool MIXED, int NT, int U>
static void run_density(Ctx& c, Gate g) {
  const size_t knot = MIXED ? sizeof(float4) : sizeof(double4);
  const size_t lds_rho = knot * (c.nr + 2);
  const PairArgs a = pair_args(c, c.cutforcesq, NT / TPR);
  const void* tab_rho = MIXED ? (const void*)c.tab32r.get() : (const void*)c.tab64.get();
  ScopedTimer tm(c, "eam_density");
  auto kd = k_eam_density<TPR, MIXED, NT, U>;
  const int grid = resident_grid(c, kd, NT, lds_rho, a.nchunks);
  kd<<<grid, NT, lds_rho, c.stream>>>(a, tab_rho, c.rdr, c.nr, c.rho.get(), g);
}

template <int TPR, bool MIXED, int NT, int U>
static void run_force(Ctx& c, int eflag, int vflag, Gate g) {
  const bool ev = eflag || vflag;
  const int E8 = NPE * c.E;
  double* gf = c.gf.get();
  double* er = ev ? c.erow.get() : nullptr;
  const size_t knot = MIXED ? sizeof(float4) : sizeof(double4);
  // mixed: the A and B planes, and C in energy/virial steps
  const size_t lds_force = (MIXED && ev ? 3 : 2) * knot * (c.nr + 2);
  const PairArgs a = pair_args(c, c.cutforcesq, NT / TPR);
  EamForceArgs q;
  q.tab = MIXED ? (const void*)c.tab32f.get() : (const void*)c.tab64.get();
  q.frho = c.frho_spl.get();
  q.rho = c.rho.get();
  q.rdr = c.rdr;
  q.rdrho = c.rdrho;
  q.rhomax = c.rhomax;
  q.nr = c.nr;
  q.nrho = c.nrho;
  q.escale_node = c.nodal_weight / 2.0;
  q.vscale_node = c.nodal_weight / 2.0;
  q.peratom = peratom_wanted(eflag, vflag) ? c.peratom.get() : nullptr;
  ScopedTimer tm(c, "eam_force", true);
  auto kf = ev ? k_eam_force<TPR, MIXED, true, NT, U> : k_eam_force<TPR, MIXED, false, NT, U>;
  const int grid = resident_grid(c, kf, NT, lds_force, a.nchunks);
  if (tm.start) {
    // timed: the events take the kernel's own start and end
    double* f = c.f.get();
    int P = c.P;
    void* args[] = {(void*)&a, (void*)&q, (void*)&f, (void*)&gf, (void*)&P, (void*)&E8, (void*)&er, (void*)&g};
    CAC_HIP(hipExtLaunchKernel(reinterpret_cast<const void*>(kf), dim3(grid), dim3(NT), args, lds_force, c.stream,
                               tm.start, tm.stop, 0));
  } else {
    kf<<<grid, NT, lds_force, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, er, g);
  }
}

// the three EAM passes; lanes per row chosen separately for the density and
// the force pass
template <bool MIXED, int NTD, int U>
static void run_density_tpr(Ctx& c, Gate g) {
  switch (c.tpr) {
    case 4: run_density<4, MIXED, NTD, U>(c, g); break;
    default: run_density<8, MIXED, NTD, U>(c, g); break;
  }
}

// The density pass stages half the table of the force pass, so more of its
// workgroups fit on a CU: its workgroup size (dens_nt) is chosen apart from
// the force pass's (NT).
template <bool MIXED, int NT, int U>
static void run_eam(Ctx& c, int eflag, int vflag, Gate g) {
  if (c.dens_nt == 1024)
    run_density_tpr<MIXED, 1024, U>(c, g);
  else
    run_density_tpr<MIXED, 512, U>(c, g);
  {
    ScopedTimer tm(c, "eam_embed");
    k_embed<MIXED><<<grid_for(c.S, 256 * RSPT), 256, 0, c.stream>>>(c.rho.get(), ucell_grid(c), c.lsdesc.get(),
                                                                    c.d_nlive.get(), c.A, c.frho_spl.get(), c.rdrho,
                                                                    c.nrho, c.site_live.get(), c.site32.get());
  }
  switch (c.tpr_force) {
    case 8: run_force<8, MIXED, NT, U>(c, eflag, vflag, g); break;
    default: run_force<4, MIXED, NT, U>(c, eflag, vflag, g); break;
  }
}

template <int TPR>
static void run_lj(Ctx& c, int eflag, int vflag, Gate g) {
  const bool ev = eflag || vflag;
  const bool mixed = c.precision == 1;
  const int E8 = NPE * c.E;
  double* er = ev ? c.erow.get() : nullptr;
  PairArgs a = pair_args(c, c.ljcutsq, PB / TPR);
  LjArgs q{c.lj1, c.lj2, c.lj3, c.lj4, c.nodal_weight / 2.0, c.nodal_weight / 2.0,
           peratom_wanted(eflag, vflag) ? c.peratom.get() : nullptr};
  ScopedTimer tm(c, "lj_force");
  auto kl = mixed ? (ev ? k_lj_force<TPR, true, true> : k_lj_force<TPR, true, false>)
                  : (ev ? k_lj_force<TPR, false, true> : k_lj_force<TPR, false, false>);
  const int grid = resident_grid(c, kl, PB, 0, a.nchunks);
  kl<<<grid, PB, 0, c.stream>>>(a, q, c.f.get(), c.gf.get(), c.P, E8, er, g);
}

static void launch_pair(Ctx& c, int eflag, int vflag, Gate g) {
  const bool mixed = c.precision == 1;
  if (c.pairstyle == 2) {
    // workgroup size and entries in flight per lane of the EAM passes
    if (mixed) {
      switch (c.pair_shape) {
        case 1: run_eam<true, 1024, 4>(c, eflag, vflag, g); break;
        case 2: run_eam<true, 512, 8>(c, eflag, vflag, g); break;
        case 3: run_eam<true, 1024, 8>(c, eflag, vflag, g); break;
        default: run_eam<true, 512, 4>(c, eflag, vflag, g); break;
      }
    } else {
      switch (c.pair_shape) {
        case 2:
        case 3: run_eam<false, 1024, 8>(c, eflag, vflag, g); break;
        default: run_eam<false, 1024, 4>(c, eflag, vflag, g); break;
      }
    }
  } else {
    switch (c.tpr_force) {
      case 8: run_lj<8>(c, eflag, vflag, g); break;
      default: run_lj<4>(c, eflag, vflag, g); break;
    }
  }
}

void ensure_inner(Ctx& c) {
  if (c.inner_mode != c.precision) {  // precision or delta changed since the last prune
    site_refresh(c, true);
    build_inner(c);
  }
}

void pair_passes(Ctx& c, int eflag, int vflag, Gate g) {
  const bool ev = eflag || vflag;
  if (ev) c.erow.reserve((size_t)7 * c.P);
  if (peratom_wanted(eflag, vflag)) c.peratom.reserve((size_t)7 * c.P);
  c.peratom_flags = ((eflag & 2) ? 1 : 0) | ((vflag & 12) ? 2 : 0);
  ensure_inner(c);
  launch_pair(c, eflag, vflag, g);
}

bool segmentable(const Ctx& c) {
  return c.precision == 1 && c.pairstyle == 2 && c.tpr == 8 && c.tpr_force == 4 && c.dens_nt == 512 &&
         c.pair_shape == 2 && !c.timing && c.P > 0;
}

int max_segments() { return MAXSEG; }

void pair_passes_segmented(Ctx* const* cs, int n, hipStream_t stream) {
  if (n < 1 || n > MAXSEG) throw HipError{hipErrorInvalidValue, "segmented pair passes: 1 to 16 models"};
  Ctx& c0 = *cs[0];
  SegRows dr{}, fr{};
  SegSites es{};
  long cd = 0, cf = 0;
  int cb = 0;
  for (int k = 0; k < n; k++) {
    Ctx& c = *cs[k];
    // the first model's tables serve all: the same potential and cutoff
    if (!segmentable(c) || c.nr != c0.nr || c.rdr != c0.rdr || c.nrho != c0.nrho || c.rdrho != c0.rdrho ||
        c.cutforcesq != c0.cutforcesq || c.inner_mode != c.precision)
      throw HipError{hipErrorInvalidValue, "segmented pair passes: models differ in potential or set-up"};
    dr.a[k] = pair_args(c, c.cutforcesq, 512 / 8);
    dr.first[k] = cd;
    cd += dr.a[k].nchunks;
    dr.out[k] = c.rho.get();
    fr.a[k] = pair_args(c, c.cutforcesq, 512 / 4);
    fr.first[k] = cf;
    cf += fr.a[k].nchunks;
    fr.out[k] = c.f.get();
    fr.gf[k] = c.gf.get();
    fr.P[k] = c.P;
    fr.E8[k] = NPE * c.E;
    es.rho[k] = c.rho.get();
    es.ug[k] = ucell_grid(c);
    es.lsdesc[k] = c.lsdesc.get();
    es.nlive[k] = c.d_nlive.get();
    es.site32[k] = c.site32.get();
    es.A[k] = c.A;
    es.first[k] = cb;
    cb += grid_for(c.S, 256 * RSPT);
  }
  dr.first[n] = cd;
  fr.first[n] = cf;
  es.first[n] = cb;
  dr.n = fr.n = es.n = n;
  const size_t lds_rho = sizeof(float4) * (c0.nr + 2);
  auto kd = k_eam_density_seg<8, 512, 8>;
  kd<<<resident_grid(c0, kd, 512, lds_rho, cd), 512, lds_rho, stream>>>(
      dr, reinterpret_cast<const float4*>(c0.tab32r.get()), c0.rdr, c0.nr);
  k_embed_seg<<<cb, 256, 0, stream>>>(es, c0.frho_spl.get(), c0.rdrho, c0.nrho);
  EamForceArgs q{};
  q.tab = c0.tab32f.get();
  q.frho = c0.frho_spl.get();
  q.rdr = c0.rdr;
  q.rdrho = c0.rdrho;
  q.rhomax = c0.rhomax;
  q.nr = c0.nr;
  q.nrho = c0.nrho;
  const size_t lds_force = 2 * sizeof(float4) * (c0.nr + 2);
  auto kf = k_eam_force_seg<4, 512, 8>;
  kf<<<resident_grid(c0, kf, 512, lds_force, cf), 512, lds_force, stream>>>(fr, q);
  CAC_HIP(hipGetLastError());
}

// eng_vdwl with ENERGY_GLOBAL, virial[6] with VIRIAL_PAIR or VIRIAL_FDOTR
// (the pair styles here never form the f.r virial: no_virial_fdotr_compute,
// pair.cpp:46, 276-279); per-point tallies alone leave both as they were
void pair_sums(Ctx& c, int eflag, int vflag) {
  if (!(eflag & 1) && !(vflag & 3)) return;
  double out[7];
  reduce_rows(c, c.erow.get(), c.P, 7, out);
  if (eflag & 1) c.eng_vdwl = out[0];
  if (vflag & 3)
    for (int i = 0; i < 6; i++) c.virial[i] = out[1 + i];
}

void pair_compute(Ctx& c, int eflag, int vflag) {
  pair_passes(c, eflag, vflag, Gate());
  if (c.E) nodal_assembly(c);
  pair_sums(c, eflag, vflag);
}

}  // namespace cac
