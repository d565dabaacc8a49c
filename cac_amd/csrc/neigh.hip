// neigh.hip -- periodic wrap, site table, counting-sort cell list and full
// per-point neighbor lists over atoms, Gauss points and interpolated atoms.
//
// Replaces, on one GPU: Domain::pbc (domain.cpp:526-738), the deformation check
// of Element::update_element_bound_box (element.cpp:1401-1465), the ghost shell of
// CommBrick::borders (comm_brick.cpp:946-1310) and Neighbor::build with
// NBinStandard / NStencilHalfBin3d / NPairHalfBinNewton (neighbor.cpp:1704-1873,
// nbin_standard.cpp:338-370, npair_half_bin_newton.cpp:25-393).
//
// The reference builds half lists with reactions; here every point (atom or
// Gauss point) gets the FULL set of sites (atoms, Gauss points and virtual atoms
// of every element, with periodic images) within rc + skin, so the force kernels
// gather and never scatter.  Periodic images are box shifts encoded in the list
// entry instead of ghost copies: a site image's position is the site position
// plus the shift, which for atoms is the same floating-point sum the reference
// forms when it creates a ghost (x + pbc*prd, atom_vec_atomic.cpp:280-345).
#include "cac_ctx.hpp"

namespace cac {


// ------------------------------------------------------------------ pbc

struct Box {
  double lo[3], hi[3], L[3];
};

static Box box_of(const Ctx& c) {
  Box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = c.lo[d];
    b.hi[d] = c.hi[d];
    b.L[d] = c.L[d];
  }
  return b;
}

// Domain::pbc on owned atoms (domain.cpp:560-640)
__global__ void k_pbc_atoms(double* __restrict__ x, int A, int P, Box b, ErrorFlags* flags) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A) return;
  for (int d = 0; d < 3; d++) {
    double xd = x[(size_t)d * P + p];
    if (!isfinite(xd)) {
      atomicOr(&flags->nan, 1);
      return;
    }
    if (xd < b.lo[d]) xd += b.L[d];
    if (xd >= b.hi[d]) {
      xd -= b.L[d];
      xd = fmax(xd, b.lo[d]);
    }
    x[(size_t)d * P + p] = xd;
  }
}

// Domain::pbc on elements: wrap the centre, shift the nodes rigidly with
// ElementVecCAC::update_node_coord(i, dim) (domain.cpp:671-737,
// element_vec_cac.cpp:3772-3784)
__global__ void k_pbc_elements(double* __restrict__ x, double* __restrict__ ex, int A, int E, int P, Box b,
                               ErrorFlags* flags) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  for (int d = 0; d < 3; d++) {
    double* nx = x + (size_t)d * P + A + (size_t)NPE * e;
    double c = ex[(size_t)d * E + e];
    bool bad = !isfinite(c);
    for (int k = 0; k < NPE; k++) bad |= !isfinite(nx[k]);
    if (bad) {
      atomicOr(&flags->nan, 1);
      return;
    }
    for (int pass = 0; pass < 2; pass++) {
      bool moved = false;
      if (pass == 0 && c < b.lo[d]) {
        c += b.L[d];
        moved = true;
      }
      if (pass == 1 && c >= b.hi[d]) {
        c -= b.L[d];
        c = fmax(c, b.lo[d]);
        moved = true;
      }
      if (moved) {
        double xt = 0.0;
        for (int k = 0; k < NPE; k++) xt += nx[k];
        xt = c - xt / NPE / 1;
        for (int k = 0; k < NPE; k++) nx[k] += xt;
      }
    }
    ex[(size_t)d * E + e] = c;
  }
}

void launch_pbc(Ctx& c) {
  Box b = box_of(c);
  if (c.A) k_pbc_atoms<<<grid_for(c.A, 256), 256, 0, c.stream>>>(c.x.get(), c.A, c.P, b, c.d_flags.get());
  if (c.E)
    k_pbc_elements<<<grid_for(c.E, 256), 256, 0, c.stream>>>(c.x.get(), c.ex.get(), c.A, c.E, c.P, b,
                                                              c.d_flags.get());
}

// Element::update_element_bound_box(check = 1): an element whose bound box grew
// or shrank more than maxelemchange = 2 in any dimension is fatal.
__global__ void k_deform(const double* __restrict__ x, double* __restrict__ init, int A, int E, int P, int first,
                         ErrorFlags* flags) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  for (int d = 0; d < 3; d++) {
    const double* nx = x + (size_t)d * P + A + (size_t)NPE * e;
    double mx = -1e30, mn = 1e30;
    for (int k = 0; k < NPE; k++) {
      mx = fmax(mx, nx[k]);
      mn = fmin(mn, nx[k]);
    }
    double size = mx - mn;
    double* ip = init + (size_t)d * E + e;
    if (first)
      *ip = size;
    else if (size / *ip > 2.0 || *ip / size > 2.0)
      atomicOr(&flags->deform, 1);
  }
}

void check_deformation(Ctx& c, bool first) {
  if (!c.E) return;
  k_deform<<<grid_for(c.E, 256), 256, 0, c.stream>>>(c.x.get(), c.initsize.get(), c.A, c.E, c.P, first ? 1 : 0,
                                                      c.d_flags.get());
}

// ------------------------------------------------------------------ sites and cells

struct Grid {
  double lo[3], L[3], cinv[3];
  int nc[3];
};

static Grid grid_of(const Ctx& c) {
  Grid g;
  for (int d = 0; d < 3; d++) {
    g.lo[d] = c.lo[d];
    g.L[d] = c.L[d];
    g.cinv[d] = c.cinv[d];
    g.nc[d] = c.nc[d];
  }
  return g;
}

// Wrap a position into the box: y = x + s*L with s in {-1,0,1}; returns the
// cell and the packed shift code (sx+1) + 3(sy+1) + 9(sz+1); -1 on a shift
// outside [-1, 1] (an element larger than the box).
__device__ __forceinline__ int wrap_cell(const Grid& g, double* y, int* code) {
  int cc[3], s[3];
  for (int d = 0; d < 3; d++) {
    double f = floor((y[d] - g.lo[d]) / g.L[d]);
    s[d] = -(int)f;
    if (s[d] < -1 || s[d] > 1) return -1;
    y[d] = y[d] + (double)s[d] * g.L[d];
    int ic = (int)floor((y[d] - g.lo[d]) * g.cinv[d]);
    cc[d] = ic < 0 ? 0 : (ic >= g.nc[d] ? g.nc[d] - 1 : ic);
  }
  *code = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
  return (cc[2] * g.nc[1] + cc[1]) * g.nc[0] + cc[0];
}

__device__ __forceinline__ void shift_of(int code, const double* L, double* sh) {
  sh[0] = (double)(code % 3 - 1) * L[0];
  sh[1] = (double)((code / 3) % 3 - 1) * L[1];
  sh[2] = (double)(code / 9 - 1) * L[2];
}

// interpolated position of unit cell u of element e (start from 0, nodes in
// order, as the reference forms Gauss-point positions, pair_eam.cpp:887-893)
__device__ __forceinline__ void interp_pos(const double* __restrict__ x, const double* __restrict__ shape, int A,
                                           int P, int e, int u, double* y) {
#pragma clang fp contract(off)
  const double* N = shape + (size_t)u * NPE;
  const size_t base = (size_t)A + (size_t)NPE * e;
  for (int d = 0; d < 3; d++) {
    const double* nx = x + (size_t)d * P + base;
    double s = 0.0;
    for (int k = 0; k < NPE; k++) s += N[k] * nx[k];
    y[d] = s;
  }
}

__global__ void k_point_occ(const double* __restrict__ x, int P, Grid g, int* __restrict__ occ,
                            ErrorFlags* flags) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double y[3] = {x[p], x[(size_t)P + p], x[2 * (size_t)P + p]};
  int code;
  int cell = wrap_cell(g, y, &code);
  if (cell < 0) {
    atomicOr(&flags->nan, 2);
    return;
  }
  occ[cell] = 1;
}

// A candidate site packed in 32 bits for the cell sort: the descriptor as an
// index (a point p as p, the interpolated atom -(1 + q) as P + q) times 32,
// plus its wrap code (0..26); the build checks P + E nucell < 2^27.
__device__ __forceinline__ unsigned pack_site(int desc, int code, int P) {
  const unsigned d = desc >= 0 ? (unsigned)desc : (unsigned)P + (unsigned)(-desc - 1);
  return (d << 5) | (unsigned)code;
}

__device__ __forceinline__ int site_desc(unsigned v, int P) {
  const int d = (int)(v >> 5);
  return d < P ? d : -(d - P + 1);
}

__global__ void k_point_sites(const double* __restrict__ x, int P, Grid g, unsigned* __restrict__ tval,
                              int* __restrict__ tcell) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double y[3] = {x[p], x[(size_t)P + p], x[2 * (size_t)P + p]};
  int code;
  int cell = wrap_cell(g, y, &code);
  if (cell < 0) {
    cell = 0;
    code = 13;
  }
  tval[p] = pack_site(p, code, P);
  tcell[p] = cell;
}

// dilate a 0/1 cell mask by +-2 cells along one axis, periodic
__global__ void k_dilate(const int* __restrict__ in, int* __restrict__ out, Grid g, int axis) {
  const long cidx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ncell = (long)g.nc[0] * g.nc[1] * g.nc[2];
  if (cidx >= ncell) return;
  int ix = (int)(cidx % g.nc[0]);
  int iy = (int)((cidx / g.nc[0]) % g.nc[1]);
  int iz = (int)(cidx / ((long)g.nc[0] * g.nc[1]));
  int v = 0;
  for (int o = -2; o <= 2; o++) {
    int c3[3] = {ix, iy, iz};
    int q = c3[axis] + o;
    q = (q < 0) ? q + g.nc[axis] : (q >= g.nc[axis] ? q - g.nc[axis] : q);
    c3[axis] = q;
    v |= in[((long)c3[2] * g.nc[1] + c3[1]) * g.nc[0] + c3[0]];
  }
  out[cidx] = v;
}

// count / fill the interpolated atoms (non-node unit cells) of each element
// whose cell lies within two cells of some point.  One workgroup per element;
// the fill pass places them with a wavefront ballot + workgroup prefix, so the
// order (element-major, unit-cell order) is deterministic.
constexpr int VB = 256;

template <bool FILL>
__global__ __launch_bounds__(VB) void k_virt(const double* __restrict__ x, const double* __restrict__ shape,
                                             const int* __restrict__ u2g, const int* __restrict__ nearmask, int A,
                                             int P, int nucell, Grid g, int* __restrict__ ecount,
                                             const int* __restrict__ eoffset,
                                             unsigned* __restrict__ tval, int* __restrict__ tcell) {
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  __shared__ int wsum[VB / 64];
  int base = FILL ? (P + eoffset[e]) : 0;
  int total = 0;
  for (int u0 = 0; u0 < nucell; u0 += VB) {
    const int u = u0 + t;
    int need = 0, cell = 0, code = 13;
    double y[3] = {0.0, 0.0, 0.0};
    if (u < nucell && u2g[u] < 0) {
      interp_pos(x, shape, A, P, e, u, y);
      cell = wrap_cell(g, y, &code);
      need = (cell >= 0) && nearmask[cell];
    }
    unsigned long long m = __ballot(need);
    int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int woff = 0, chunk = 0;
    for (int w = 0; w < VB / 64; w++) {
      if (w < wave) woff += wsum[w];
      chunk += wsum[w];
    }
    if (FILL && need) {
      const int s = base + total + woff + rank;
      tval[s] = pack_site(-(1 + e * nucell + u), code, P);
      tcell[s] = cell;
    }
    total += chunk;
    __syncthreads();
  }
  if (!FILL && t == 0) ecount[e] = total;
}

// sorted (cell, descriptor) pairs -> site descriptors, packed cell coordinates, row flags
__global__ void k_unpack_sites(const int* __restrict__ key, const unsigned* __restrict__ val, int S, int P, Grid g,
                               int* __restrict__ sdesc, int* __restrict__ simg, int* __restrict__ scell,
                               int* __restrict__ isrow) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S) return;
  const int cid = key[k];
  const unsigned v = val[k];
  const int desc = site_desc(v, P);
  scell[k] = cell_pack(cid % g.nc[0], (cid / g.nc[0]) % g.nc[1], cid / (g.nc[0] * g.nc[1]));
  sdesc[k] = desc;
  simg[k] = (int)(v & 31u);
  isrow[k] = desc >= 0;
}

// rows in site order: rowidx is the exclusive scan of isrow; a count other
// than P is flagged (read with the list build's status) and no row past P is
// written
__global__ void k_rows(const int* __restrict__ sdesc, const int* __restrict__ isrow, const int* __restrict__ rowidx,
                       int S, int P, int* __restrict__ row_point, int* __restrict__ row_site, ErrorFlags* flags) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S) return;
  if (k == S - 1 && rowidx[k] + isrow[k] != P) flags->rows = 1;
  if (!isrow[k]) return;
  const int r = rowidx[k];
  if (r >= P) return;
  row_point[r] = sdesc[k];
  row_site[r] = k;
}

// The reference's form of a squared pair distance (npair_half_bin_newton.cpp:
// 89-141, 194-199, 284-289, 346-364), for a pair whose squared distance lies
// within a relative 1e-9 of the list cutoff; the fast form of the list scan
// decides every other pair, where the two cannot disagree.
//
// The reference subtracts from a row point's own position its neighbor's
// ghost coordinate: atom x + T L, or for an interpolated atom N_k (x_k + T L)
// node by node, T L being the box shift of the ghost copy
// (atom_vec_atomic.cpp:280-345, element_vec_cac.cpp:596-821).  Gauss points
// sit on nodes (N is 0 or 1), so a Gauss point's position is its node's
// coordinate exactly, as a row and as a neighbor.  The site records hold each
// site wrapped into the box by its own shift W (an interpolated atom as
// sum_k N_k x_k + W L), so the pair's ghost shift is T = W_site + t_entry -
// W_row, and the fast form rounds differently.
//
// A pair of two points is tested once in the reference, from the row that
// holds it in the half list: an atom-Gauss pair from the atom's row (Gauss
// rows hold no atoms); a Gauss-Gauss pair from the lower element index, or
// the lower Gauss cell within one element, and across the periodic boundary
// from the element whose neighbor's ghost centre lies above-right of its own
// (z, then y, then x; npair_half_bin_newton.cpp:346-357).  An atom-atom pair
// follows the reference's bins, which this build does not form: it is tested
// from the row at hand (with T = 0 both directions round alike).
constexpr double NEAR_CUT_REL = 1e-9;

struct RefOrder {
  const double* ex;  // element centres [3][E], wrapped as Domain::pbc leaves them
  int E;
  int n2g[8];        // node -> Gauss cell
};

// does the reference store the Gauss-Gauss pair (element ei cell gi, element
// ej cell gj seen through ghost shift T) in ei's row?
__device__ __forceinline__ bool gauss_pair_in_row(const RefOrder& ro, const double* L, int ei, int gi, int ej, int gj,
                                                  const int* T) {
  if (T[0] == 0 && T[1] == 0 && T[2] == 0) return ei < ej || (ei == ej && gi < gj);
  double ci[3], cj[3];
  for (int d = 0; d < 3; d++) {
    ci[d] = ro.ex[(size_t)d * ro.E + ei];
    cj[d] = ro.ex[(size_t)d * ro.E + ej] + (double)T[d] * L[d];
  }
  if (cj[2] < ci[2]) return false;
  if (cj[2] == ci[2]) {
    if (cj[1] < ci[1]) return false;
    if (cj[1] == ci[1] && cj[0] < ci[0]) return false;
  }
  return true;
}

__device__ __forceinline__ double reference_rsq(const double* __restrict__ x, int P, int A, const UcellGrid& ug,
                                             const RefOrder& ro, const double* L, int prow, int wrow, int dsite,
                                             int wsite, unsigned tcode) {
#pragma clang fp contract(off)
  int T[3];
  T[0] = (wsite % 3 - 1) + ((int)(tcode % 3u) - 1) - (wrow % 3 - 1);
  T[1] = ((wsite / 3) % 3 - 1) + ((int)((tcode / 3u) % 3u) - 1) - ((wrow / 3) % 3 - 1);
  T[2] = (wsite / 9 - 1) + ((int)(tcode / 9u) - 1) - (wrow / 9 - 1);
  double del[3];
  if (dsite >= 0) {
    // two points: test from the row the reference keeps the pair in
    int from = prow, to = dsite, sgn = 1;
    if (prow >= A && dsite < A) {
      from = dsite;  // atom-Gauss: the atom's row
      to = prow;
      sgn = -1;
    } else if (prow >= A && dsite >= A) {
      const int ei = (prow - A) / NPE, ej = (dsite - A) / NPE;
      const int gi = ro.n2g[prow - A - NPE * ei], gj = ro.n2g[dsite - A - NPE * ej];
      if (!gauss_pair_in_row(ro, L, ei, gi, ej, gj, T)) {
        from = dsite;
        to = prow;
        sgn = -1;
      }
    }
    for (int d = 0; d < 3; d++)
      del[d] = x[(size_t)d * P + from] - (x[(size_t)d * P + to] + (double)(sgn * T[d]) * L[d]);
  } else {
    int e, u;
    split_ucell(ug, -dsite - 1, e, u);
    double N[NPE];
    hex_shape_at(ug, u, N);
    const size_t base = (size_t)A + (size_t)NPE * e;
    for (int d = 0; d < 3; d++) {
      double v = x[(size_t)d * P + prow];
      const double sh = (double)T[d] * L[d];
      for (int k = 0; k < NPE; k++) v -= N[k] * (x[(size_t)d * P + base + k] + sh);
      del[d] = v;
    }
  }
  return del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
}

// BL lanes per row scan the 5x5x5 block of cells around the row's cell (cell
// edge > (rc+skin)/2, periodic) and keep every site image within rc + skin
// except the row's own site.  Sites are sorted by cell, so the five cells of
// one x strip are one contiguous site range (two or three across the periodic
// seams): the lanes read it coalesced, and a ballot + prefix per lane group
// appends the kept sites in scan order (z, y, x cell order, site order inside
// a cell), so the list is deterministic.  A box of two to four cells along an
// axis (rc + skin <= L < 2.5 (rc + skin)) meets a cell more than once in the
// block, each time as another image, as the reference's ghost shell holds
// several images of a site when the box is short.  Only an overflowing row
// touches the shared status words.
constexpr int BL = 8;

//
// The scan runs twice.  The first pass (REF false) keeps a candidate by the
// record difference alone and marks, in nearrow, each row holding a candidate
// within NEAR_CUT_REL of the cutoff; the second (REF true) visits only the
// marked rows and scans them again from the start, deciding the near
// candidates by reference_rsq.  The rare recompute then costs the common scan
// no registers.
template <bool REF>
__global__ __launch_bounds__(256) void k_build_lists(const double4* __restrict__ site, const int* __restrict__ row_site,
                                                     int R, Grid g, const int* __restrict__ cstart, double cutsq,
                                                     int* __restrict__ nnum, unsigned* __restrict__ nlist, int cap,
                                                     ErrorFlags* flags, int* __restrict__ nearrow,
                                                     const double* __restrict__ x, int P, int A, UcellGrid ug,
                                                     RefOrder ro, const int* __restrict__ sdesc,
                                                     const int* __restrict__ simg) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x % BL;
  const int gbase = (threadIdx.x & 63) & ~(BL - 1);
  const int r = (int)(((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / BL);
  if (r >= R) return;
  if (REF && !nearrow[r]) return;
  bool nearhit = false;
  const int k = row_site[r];
  const double4 pk = site[k];
  const double px = pk.x, py = pk.y, pz = pk.z;
  const double near = NEAR_CUT_REL * cutsq;
  int ic[3];
  const double pp[3] = {px, py, pz};
  for (int d = 0; d < 3; d++) {
    int v = (int)floor((pp[d] - g.lo[d]) * g.cinv[d]);
    ic[d] = v < 0 ? 0 : (v >= g.nc[d] ? g.nc[d] - 1 : v);
  }
  unsigned* row = nlist + (size_t)r * cap;
  const int nx = g.nc[0];
  const int xlo = ic[0] - 2, xhi = ic[0] + 2;
  int n = 0;
  for (int dz = -2; dz <= 2; dz++) {
    int cz = ic[2] + dz, tz = 0;
    if (cz < 0) { cz += g.nc[2]; tz = -1; }
    else if (cz >= g.nc[2]) { cz -= g.nc[2]; tz = 1; }
    const double shz = (double)tz * g.L[2];
    for (int dy = -2; dy <= 2; dy++) {
      int cy = ic[1] + dy, ty = 0;
      if (cy < 0) { cy += g.nc[1]; ty = -1; }
      else if (cy >= g.nc[1]) { cy -= g.nc[1]; ty = 1; }
      const double shy = (double)ty * g.L[1];
      const long cbase = ((long)cz * g.nc[1] + cy) * nx;
      for (int tx = -1; tx <= 1; tx++) {
        // the strip's cells in dx order: the part across the -x seam, the part
        // inside the box, the part across the +x seam (a box of fewer than
        // five cells holds all three, and a cell may come twice, as two images)
        int a, b;
        if (tx < 0) {
          if (xlo >= 0) continue;
          a = xlo + nx;
          b = nx - 1;
        } else if (tx > 0) {
          if (xhi < nx) continue;
          a = 0;
          b = xhi - nx;
        } else {
          a = xlo > 0 ? xlo : 0;
          b = xhi < nx ? xhi : nx - 1;
        }
        const double shx = (double)tx * g.L[0];
        const unsigned tcode = (unsigned)((tx + 1) + 3 * (ty + 1) + 9 * (tz + 1));
        const int j0 = cstart[cbase + a], j1 = cstart[cbase + b + 1];
        for (int jb = j0; jb < j1; jb += BL) {
          const int j = jb + lane;
          bool keep = false;
          if (j < j1) {
            const double4 sj = site[j];
            const double delx = px - (sj.x + shx);
            const double dely = py - (sj.y + shy);
            const double delz = pz - (sj.z + shz);
            double rsq = delx * delx + dely * dely + delz * delz;
            if (fabs(rsq - cutsq) <= near) {
              if (REF) rsq = reference_rsq(x, P, A, ug, ro, g.L, sdesc[k], simg[k], sdesc[j], simg[j], tcode);
              else nearhit = true;
            }
            keep = rsq <= cutsq && !(j == k && tcode == 13u);
          }
          const unsigned long long m = __ballot(keep);
          const unsigned gm = (unsigned)((m >> gbase) & ((1ull << BL) - 1ull));
          const int rank = __popc(gm & ((1u << lane) - 1u));
          if (keep && n + rank < cap) row[n + rank] = (unsigned)j | (tcode << IMG_SHIFT);
          n += __popc(gm);
        }
      }
    }
  }
  if (!REF) {
    const unsigned long long h = __ballot(nearhit);
    if (lane == 0) nearrow[r] = ((h >> gbase) & ((1ull << BL) - 1ull)) != 0ull;
  }
  if (lane == 0) {
    nnum[r] = n < cap ? n : cap;
    if (n > cap) {
      atomicMax(&flags->maxcount, n);
      atomicOr(&flags->overflow, 1);
    }
  }
}

// keep only the sites some row references (and every row's own site);
// 16 lanes per row read the row coalesced
__global__ __launch_bounds__(256) void k_mark_used(const int* __restrict__ row_site, const int* __restrict__ nnum,
                                                   const unsigned* __restrict__ nlist, int R, int cap,
                                                   int* __restrict__ used) {
  const int lane = threadIdx.x & 15;
  const int r = (int)(((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 4);
  if (r >= R) return;
  if (lane == 0) used[row_site[r]] = 1;
  const unsigned* row = nlist + (size_t)r * cap;
  const int n = nnum[r];
  for (int i = lane; i < n; i += 16) used[row[i] & SITE_MASK] = 1;
}

// the descriptors of the kept sites; their records are written by the full
// refresh that follows the build
__global__ void k_compact_sites(const int* __restrict__ used, const int* __restrict__ newidx, int S,
                                const int* __restrict__ sdesc, const int* __restrict__ simg,
                                const int* __restrict__ scell, int* __restrict__ sdesc2, int* __restrict__ simg2,
                                int* __restrict__ scell2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S || !used[k]) return;
  const int d = newidx[k];
  sdesc2[d] = sdesc[k];
  simg2[d] = simg[k];
  scell2[d] = scell[k];
}

__global__ __launch_bounds__(256) void k_remap_rows(int* __restrict__ row_site, const int* __restrict__ nnum,
                                                    unsigned* __restrict__ nlist, int R, int cap,
                                                    const int* __restrict__ newidx) {
  const int lane = threadIdx.x & 15;
  const int r = (int)(((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 4);
  if (r >= R) return;
  if (lane == 0) row_site[r] = newidx[row_site[r]];
  unsigned* row = nlist + (size_t)r * cap;
  const int n = nnum[r];
  for (int i = lane; i < n; i += 16) {
    const unsigned e = row[i];
    row[i] = (unsigned)newidx[e & SITE_MASK] | (e & ~SITE_MASK);
  }
}

static void read_flags(Ctx& c) {
  static_assert(sizeof(ErrorFlags) == 8 * sizeof(int), "ErrorFlags is eight ints");
  read_device_ints(c, reinterpret_cast<const int*>(c.d_flags.get()), 8, reinterpret_cast<int*>(c.h_flags));
}

// Domain::pbc on the model alone, as write_data calls it before writing
// (write_data.cpp:120-131): element centres from their nodes, then the wrap.
// Points move by box lengths, so the lists no longer match the positions: the
// caller marks them stale and the next setup builds them again.
void wrap_model(Ctx& c) {
  CAC_HIP(hipMemsetAsync(c.d_flags.get(), 0, sizeof(ErrorFlags), c.stream));
  if (c.E) update_centers(c);
  launch_pbc(c);
  read_flags(c);
  if (c.h_flags->nan & 1) throw HipError{hipErrorInvalidValue, "Non-numeric positions - simulation unstable"};
}

void build_neighbor(Ctx& c) {
  ScopedTimer tm(c, "neigh_build");
  c.ago = 0;
  c.ncalls++;
  CAC_HIP(hipMemsetAsync(c.d_flags.get(), 0, sizeof(ErrorFlags), c.stream));
  // step words of an earlier speculative step are spent
  CAC_HIP(hipMemsetAsync(c.d_steps.get(), 0xFF, sizeof(int) * 2, c.stream));
  c.h_steps[0] = c.h_steps[1] = -1;
  if (c.E) update_centers(c);
  launch_pbc(c);
  check_deformation(c, false);
  // hold positions for check_distance (neighbor.cpp:1732-1776)
  CAC_HIP(hipMemcpyAsync(c.xhold.get(), c.x.get(), sizeof(double) * 3 * c.P, hipMemcpyDeviceToDevice, c.stream));

  // cell grid: edge > (rc + skin)/2 so a 5x5x5 block covers the cutoff
  for (int d = 0; d < 3; d++) {
    int n = (int)floor(c.L[d] / (0.5 * c.cutneigh * (1.0 + 1e-9)));
    if (n < 2) throw HipError{hipErrorInvalidValue, "box shorter than rc + skin in some dimension"};
    c.nc[d] = n;
    c.cinv[d] = (double)n / c.L[d];
  }
  c.ncells = (long)c.nc[0] * c.nc[1] * c.nc[2];
  if (c.nc[0] >= (1 << CELL_BITS_X) || c.nc[1] >= (1 << CELL_BITS_Y) || c.nc[2] >= (1 << CELL_BITS_Z))
    throw HipError{hipErrorInvalidValue, "too many neighbor cells along a box edge"};
  Grid g = grid_of(c);
  const int P = c.P;

  c.occ.reserve(c.ncells);
  c.nearA.reserve(c.ncells);
  c.nearB.reserve(c.ncells);
  CAC_HIP(hipMemsetAsync(c.occ.get(), 0, sizeof(int) * c.ncells, c.stream));

  // pass 1: cells holding a point, dilated by two cells
  k_point_occ<<<grid_for(P, 256), 256, 0, c.stream>>>(c.x.get(), P, g, c.occ.get(), c.d_flags.get());
  const int cb = grid_for(c.ncells, 256);
  k_dilate<<<cb, 256, 0, c.stream>>>(c.occ.get(), c.nearA.get(), g, 0);
  k_dilate<<<cb, 256, 0, c.stream>>>(c.nearA.get(), c.nearB.get(), g, 1);
  k_dilate<<<cb, 256, 0, c.stream>>>(c.nearB.get(), c.nearA.get(), g, 2);

  // pass 2: count interpolated atoms near any point, then lay out the sites
  int nvirt = 0;
  if (c.E) {
    c.ecount.reserve(c.E);
    c.eoffset.reserve(c.E);
    k_virt<false><<<c.E, VB, 0, c.stream>>>(c.x.get(), c.shape.get(), c.u2g.get(), c.nearA.get(), c.A, P, c.nucell,
                                           g, c.ecount.get(), nullptr, nullptr, nullptr);
    scan_exclusive(c, c.ecount.get(), c.eoffset.get(), c.E, &nvirt);
  }
  const int S = P + nvirt;
  if ((long)S >= (1L << IMG_SHIFT)) throw HipError{hipErrorInvalidValue, "more than 2^27 sites"};
  // the cell sort packs a candidate's descriptor index in 27 bits (pack_site)
  if ((long)P + (long)c.E * (long)c.nucell >= (1L << 27))
    throw HipError{hipErrorInvalidValue, "more than 2^27 atoms, Gauss points and interpolated atoms"};
  c.tval.reserve(S);
  c.tcell.reserve(S);
  k_point_sites<<<grid_for(P, 256), 256, 0, c.stream>>>(c.x.get(), P, g, c.tval.get(), c.tcell.get());
  if (c.E && nvirt)
    k_virt<true><<<c.E, VB, 0, c.stream>>>(c.x.get(), c.shape.get(), c.u2g.get(), c.nearA.get(), c.A, P, c.nucell, g,
                                          nullptr, c.eoffset.get(), c.tval.get(), c.tcell.get());

  // sites ordered by cell (stable: candidate order inside a cell), then their
  // FP64 records from the descriptors, with the refresh's arithmetic
  sort_sites_by_cell(c, c.tcell.get(), c.tval.get(), S);
  c.site.reserve(S);
  c.sdesc.reserve(S);
  c.simg.reserve(S);
  c.isflag.reserve(S);
  c.row_point.reserve(P);
  c.row_site.reserve(P);
  c.scell.reserve(S);
  k_unpack_sites<<<grid_for(S, 256), 256, 0, c.stream>>>(c.sort_keys.get(), c.sort_vals.get(), S, P, g, c.sdesc.get(),
                                                         c.simg.get(), c.scell.get(), c.isflag.get());
  c.S = S;
  site_positions(c);
  c.rowidx.reserve(S);
  scan_exclusive(c, c.isflag.get(), c.rowidx.get(), S, nullptr);
  k_rows<<<grid_for(S, 256), 256, 0, c.stream>>>(c.sdesc.get(), c.isflag.get(), c.rowidx.get(), S, P,
                                                 c.row_point.get(), c.row_site.get(), c.d_flags.get());
  order_rows(c);
  c.nvirt = nvirt;

  // full lists; grow the per-point capacity and rebuild once if a row overflowed
  if (c.cap == 0) {  // first build: the mean count at the site density, with margin
    const double vol = c.L[0] * c.L[1] * c.L[2];
    const double est = (double)S / vol * 4.18879020478639 * c.cutneighsq * c.cutneigh;
    c.cap = std::min(c.one, ((int)(1.5 * est) + 32 + 15) / 16 * 16);
  }
  c.nnum.reserve(P);
  c.nearrow.reserve(P);
  RefOrder ro;
  ro.ex = c.E ? c.ex.get() : nullptr;
  ro.E = c.E;
  for (int k = 0; k < 8; k++) ro.n2g[k] = c.n2g[k];
  for (int attempt = 0; attempt < 3; attempt++) {
    c.nlist.reserve((size_t)c.cap * P);
    CAC_HIP(hipMemsetAsync(&c.d_flags.get()->overflow, 0, sizeof(int) * 2, c.stream));
    const UcellGrid ug = ucell_grid(c);
    unsigned* nl = reinterpret_cast<unsigned*>(c.nlist.get());
    k_build_lists<false><<<grid_for((long)BL * P, 256), 256, 0, c.stream>>>(
        c.site.get(), c.row_site.get(), P, g, c.cell_start.get(), c.cutneighsq, c.nnum.get(), nl, c.cap,
        c.d_flags.get(), c.nearrow.get(), c.x.get(), P, c.A, ug, ro, c.sdesc.get(), c.simg.get());
    k_build_lists<true><<<grid_for((long)BL * P, 256), 256, 0, c.stream>>>(
        c.site.get(), c.row_site.get(), P, g, c.cell_start.get(), c.cutneighsq, c.nnum.get(), nl, c.cap,
        c.d_flags.get(), c.nearrow.get(), c.x.get(), P, c.A, ug, ro, c.sdesc.get(), c.simg.get());
    read_flags(c);
    if (!c.h_flags->overflow) break;
    int need = c.h_flags->maxcount;
    if (need > c.one) throw HipError{hipErrorInvalidValue, "Neighbor list overflow, boost neigh_modify one"};
    c.cap = ((need + 15) / 16) * 16 + 16;
  }
  if (c.h_flags->rows) throw HipError{hipErrorUnknown, "row count mismatch"};
  if (c.h_flags->nan & 1) throw HipError{hipErrorInvalidValue, "Non-numeric positions - simulation unstable"};
  if (c.h_flags->nan & 2) throw HipError{hipErrorInvalidValue, "an element is larger than the periodic box"};
  if (c.h_flags->deform) throw HipError{hipErrorInvalidValue, "Simulation unstable! an element deformed by more than 2.0 times"};
  // drop the sites no list references: the per-step refresh and the gathers
  // then touch only live sites (order, and so cell locality, is kept)
  {
    c.used.reserve(S);
    c.newidx.reserve(S);
    CAC_HIP(hipMemsetAsync(c.used.get(), 0, sizeof(int) * S, c.stream));
    unsigned* nl = reinterpret_cast<unsigned*>(c.nlist.get());
    k_mark_used<<<grid_for(16L * P, 256), 256, 0, c.stream>>>(c.row_site.get(), c.nnum.get(), nl, P, c.cap,
                                                               c.used.get());
    int S2 = 0;
    scan_exclusive(c, c.used.get(), c.newidx.get(), S, &S2);
    c.sdesc2.reserve(S2);
    c.simg2.reserve(S2);
    c.scell2.reserve(S2);
    k_compact_sites<<<grid_for(S, 256), 256, 0, c.stream>>>(c.used.get(), c.newidx.get(), S, c.sdesc.get(),
                                                            c.simg.get(), c.scell.get(), c.sdesc2.get(),
                                                            c.simg2.get(), c.scell2.get());
    k_remap_rows<<<grid_for(16L * P, 256), 256, 0, c.stream>>>(c.row_site.get(), c.nnum.get(), nl, P, c.cap,
                                                                c.newidx.get());
    c.sdesc.swap(c.sdesc2);
    c.simg.swap(c.simg2);
    c.scell.swap(c.scell2);
    c.S = S2;
    c.nvirt = S2 - P;
  }

  c.entries = -1;  // stored entries, counted on request
  c.built = true;
  c.inner_mode = -1;
  site_refresh(c, true);  // fills the mixed path's cell-relative copy
  build_inner(c);
}

// ------------------------------------------------------------------ check_distance (neighbor.cpp:1617-1683)
// One pass tests both lists: the reference's trigger (any point moved more
// than skin/2 since the last build) and the inner list's (moved more than
// delta/2 since it was pruned).  A virtual atom moves by a convex combination
// of its element's node moves, so points bound every site.

__global__ void k_check_distance(const double* __restrict__ x, const double* __restrict__ xh,
                                 const double* __restrict__ xhi, int P, double trig, double trig_in, int outer,
                                 ErrorFlags* flags) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int hit = 0, hit_in = 0;
  if (p < P) {
    const double x0 = x[p], x1 = x[(size_t)P + p], x2 = x[2 * (size_t)P + p];
    if (outer) {
      const double dx = x0 - xh[p], dy = x1 - xh[(size_t)P + p], dz = x2 - xh[2 * (size_t)P + p];
      hit = (dx * dx + dy * dy + dz * dz > trig);
    }
    const double ex = x0 - xhi[p], ey = x1 - xhi[(size_t)P + p], ez = x2 - xhi[2 * (size_t)P + p];
    hit_in = (ex * ex + ey * ey + ez * ez > trig_in);
  }
  const bool lead = (threadIdx.x & 63) == 0;
  if (__ballot(hit) && lead) atomicOr(&flags->rebuild, 1);
  if (__ballot(hit_in) && lead) atomicOr(&flags->rebuild_in, 1);
}

__global__ void k_clear_checks(ErrorFlags* flags) {
  flags->rebuild = 0;
  flags->rebuild_in = 0;
}

int check_distance(Ctx& c, bool outer, int* inner) {
  k_clear_checks<<<1, 1, 0, c.stream>>>(c.d_flags.get());
  const double hd = 0.5 * c.inner_skin;
  k_check_distance<<<grid_for(c.P, 256), 256, 0, c.stream>>>(c.x.get(), c.xhold.get(), c.xhold_in.get(), c.P,
                                                             c.triggersq, hd * hd, outer ? 1 : 0, c.d_flags.get());
  CAC_HIP(hipMemcpyAsync(&c.h_flags->rebuild, &c.d_flags.get()->rebuild, 2 * sizeof(int), hipMemcpyDeviceToHost,
                         c.stream));
  CAC_HIP(hipStreamSynchronize(c.stream));
  if (inner) *inner = c.h_flags->rebuild_in;
  return c.h_flags->rebuild;
}

// ------------------------------------------------------------------ inner list
// 16 lanes per row walk the outer list and keep the entries within rc + delta;
// a ballot per 16-lane group and a prefix of its bits place them in outer-list
// order, so the inner list is deterministic.  It has the outer list's row
// capacity, so it cannot overflow and needs no status check.  The mixed path
// re-encodes each entry as (site | cell offset << 25): the site image sits
// (dx, dy, dz) cells from the row's cell, which with cell-relative FP32
// coordinates gives the pair vector as (u_i - u_j) - d * cell edge.

constexpr int MIX_SHIFT = 25;

// Each row also holds its point's position for the next inner-list check.
// The mixed path tests distances on the FP32 cell-relative records of every
// site, with the cutoff widened by 1e-5 relative: that covers FP32 rounding
// of the pair vector (below 1e-6 relative at these lengths), so the inner
// list still holds every pair within rc + delta.  The FP64 path tests the FP64
// records.  A mixed record carries its site's packed cell in w, so one gather
// per entry gives both the position and the cell.
//
// 16 lanes per row, PU entries per lane per batch (PU = 9: one batch holds the
// 134-entry rows of an FCC crystal at rc + skin).  List words are read at an
// index clamped into the row and masked past its end: no branch.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_prune(const double4* __restrict__ site, const float4* __restrict__ site32f,
                                               const int* __restrict__ row_site, const int* __restrict__ row_point,
                                               const int* __restrict__ nnum, const unsigned* __restrict__ nlist,
                                               int R, int cap, Grid g, double cutsq_in, int* __restrict__ innum,
                                               unsigned* __restrict__ ilist, int icap, const double* __restrict__ x,
                                               double* __restrict__ xhold_in, unsigned char* __restrict__ live) {
  constexpr int PU = MIXED ? 9 : 6;
  // LDS tables of the mixed test: the cell shift of each of the 27 image codes,
  // packed as the cell-index difference it adds, and the 125 cell-offset
  // vectors (-2..2 cells per axis) times the cell edge
  __shared__ int img_cells[27][3];
  __shared__ float4 off_vec[128];
  if (MIXED) {
    for (int t = threadIdx.x; t < 27; t += blockDim.x) {
      img_cells[t][0] = (t % 3 - 1) * g.nc[0];
      img_cells[t][1] = (t / 3 % 3 - 1) * g.nc[1];
      img_cells[t][2] = (t / 9 - 1) * g.nc[2];
    }
    for (int t = threadIdx.x; t < 125; t += blockDim.x)
      off_vec[t] = make_float4((float)(t % 5 - 2) * (float)(g.L[0] / g.nc[0]),
                               (float)(t / 5 % 5 - 2) * (float)(g.L[1] / g.nc[1]),
                               (float)(t / 25 - 2) * (float)(g.L[2] / g.nc[2]), 0.0f);
    __syncthreads();
  }
  const int lane = threadIdx.x & 15;
  const int gbase = (threadIdx.x & 63) & ~15;
  // grid-stride over rows, so a step that does not re-prune launches few workgroups
  const long rstride = ((long)gridDim.x * blockDim.x) >> 4;
  const long rfirst = ((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 4;
  const float cut32 = (float)(cutsq_in * (1.0 + 1e-5));
  for (long rl = rfirst; rl < R; rl += rstride) {
    const int r = (int)rl;
    const int k = row_site[r];
    if (lane < 3) {
      const size_t i = (size_t)lane * R + row_point[r];
      xhold_in[i] = x[i];
    }
    if (lane == 3) live[k] = 1;
    const int n = nnum[r];
    const double4 pk = MIXED ? make_double4(0.0, 0.0, 0.0, 0.0) : site[k];
    const float4 pk32 = MIXED ? site32f[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int cp = MIXED ? __float_as_int(pk32.w) : 0;
    const int icx = cell_ix(cp), icy = cell_iy(cp), icz = cell_iz(cp);
    const unsigned* row = nlist + (size_t)r * cap;
    unsigned* out = ilist + (size_t)r * icap;
    int count = 0;
    for (int base = 0; base < n; base += 16 * PU) {
      unsigned ent[PU];
#pragma unroll
      for (int u = 0; u < PU; u++) {
        const int jj = base + 16 * u + lane;
        ent[u] = row[jj < n ? jj : n - 1];
      }
      double4 sj[MIXED ? 1 : PU];
      float4 sj32[MIXED ? PU : 1];
#pragma unroll
      for (int u = 0; u < PU; u++) {
        if (MIXED)
          sj32[u] = site32f[ent[u] & SITE_MASK];
        else
          sj[u] = site[ent[u] & SITE_MASK];
      }
#pragma unroll
      for (int u = 0; u < PU; u++) {
        const int jj = base + 16 * u + lane;
        const unsigned t = ent[u] >> IMG_SHIFT;
        bool keep;
        unsigned eout = ent[u];
        if (MIXED) {
          // the site image's cell relative to the row's cell, in -2..2 per axis
          const int cj = __float_as_int(sj32[u].w);
          const int ox = cell_ix(cj) + img_cells[t][0] - icx;
          const int oy = cell_iy(cj) + img_cells[t][1] - icy;
          const int oz = cell_iz(cj) + img_cells[t][2] - icz;
          const unsigned dcode = (unsigned)(ox + 5 * oy + 25 * oz + 62);
          const float4 o = off_vec[dcode];
          const float dx = (pk32.x - sj32[u].x) - o.x;
          const float dy = (pk32.y - sj32[u].y) - o.y;
          const float dz = (pk32.z - sj32[u].z) - o.z;
          keep = jj < n && dx * dx + dy * dy + dz * dz <= cut32;
          eout = (ent[u] & SITE_MASK) | (dcode << MIX_SHIFT);
        } else {
          const int tx = (int)(t % 3u) - 1, ty = (int)((t / 3u) % 3u) - 1, tz = (int)(t / 9u) - 1;
          const double dx = pk.x - (sj[u].x + (double)tx * g.L[0]);
          const double dy = pk.y - (sj[u].y + (double)ty * g.L[1]);
          const double dz = pk.z - (sj[u].z + (double)tz * g.L[2]);
          keep = jj < n && dx * dx + dy * dy + dz * dz <= cutsq_in;
        }
        const unsigned long long m = __ballot(keep);
        const unsigned gm = (unsigned)((m >> gbase) & 0xFFFFull);
        const int rank = __popc(gm & ((1u << lane) - 1u));
        if (keep) out[count + rank] = eout;
        // a point's site is flagged by the point's own row (above); only the
        // other sites are flagged here, which in the mixed path skips the
        // scattered byte stores of atom-atom pairs
        if (keep && !(MIXED && (__float_as_uint(sj32[u].w) & POINT_SITE_BIT))) live[ent[u] & SITE_MASK] = 1;
        count += __popc(gm);
      }
    }
    if (lane == 0) innum[r] = count;
  }
}

// dense index of every row's site, and the inner-list entries re-pointed at
// the dense records (16 lanes per row; a lane reads all its entries, then all
// their dense indices, then writes)

// Each row is then padded to a multiple of ROW_PAD entries (the pair passes
// read whole batches) with an entry naming the dummy record at index nlive,
// far from every point, which the first thread writes: the pair passes find
// it beyond every cutoff.
constexpr double DUMMY_COORD = 1e5;  // dummy record coordinates (A or sigma)

__global__ __launch_bounds__(256) void k_live_remap(const int* __restrict__ idx, const int* __restrict__ row_site,
                                                    int* __restrict__ row_live, const int* __restrict__ innum,
                                                    unsigned* __restrict__ ilist, int R, int cap, unsigned mask,
                                                    unsigned pad_code, const int* __restrict__ nlive,
                                                    float4* __restrict__ site32, double4* __restrict__ site_live) {
  constexpr int RU = 6;
  const int lane = threadIdx.x & 15;
  const unsigned dummy = (unsigned)*nlive | pad_code;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (site32) site32[*nlive] = make_float4((float)DUMMY_COORD, (float)DUMMY_COORD, (float)DUMMY_COORD, 0.0f);
    if (site_live) site_live[*nlive] = make_double4(DUMMY_COORD, DUMMY_COORD, DUMMY_COORD, 0.0);
  }
  const long rstride = ((long)gridDim.x * blockDim.x) >> 4;
  const long rfirst = ((long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 4;
  for (long rl = rfirst; rl < R; rl += rstride) {
    const int r = (int)rl;
    if (lane == 0) row_live[r] = idx[row_site[r]];
    unsigned* row = ilist + (size_t)r * cap;
    const int n = innum[r];
    for (int b = 0; b < n; b += 16 * RU) {
      unsigned e[RU];
      int d[RU];
#pragma unroll
      for (int u = 0; u < RU; u++) {
        const int i = b + 16 * u + lane;
        e[u] = row[i < n ? i : n - 1];
      }
#pragma unroll
      for (int u = 0; u < RU; u++) d[u] = idx[e[u] & mask];
#pragma unroll
      for (int u = 0; u < RU; u++) {
        const int i = b + 16 * u + lane;
        if (i < n) row[i] = (unsigned)d[u] | (e[u] & ~mask);
      }
    }
    const int np = (n + ROW_PAD - 1) / ROW_PAD * ROW_PAD;  // <= icap, a multiple of 16 >= n
    if (n + lane < np) row[n + lane] = dummy;
  }
}

void build_inner(Ctx& c) {
  ScopedTimer tm(c, "inner_build");
  const bool mixed = c.precision == 1;
  if (mixed && (long)c.S >= (1L << MIX_SHIFT)) throw HipError{hipErrorInvalidValue, "more than 2^25 sites for the mixed list"};
  const double rin = c.cutforce + c.inner_skin;
  c.inner_cutsq = rin * rin;
  Grid g = grid_of(c);
  c.icap = (c.cap + 15) / 16 * 16;  // a row padded to a multiple of ROW_PAD entries fits
  // the pair passes address the inner list with 32-bit byte offsets
  if ((size_t)c.P * (size_t)c.icap * 4 >= (1ull << 32))
    throw HipError{hipErrorInvalidValue, "inner neighbor list larger than 4 GiB (too many points for one GPU)"};
  c.innum.reserve(c.P);
  c.ilist.reserve((size_t)c.icap * c.P);
  c.live_flag.reserve(c.S);
  c.live_idx.reserve(c.S);
  c.row_live.reserve(c.P);
  c.d_nlive.reserve(1);
  // the last compaction left its flags zero; a new allocation or more sites
  // than it read are cleared here
  if (c.live_flag.get() != c.live_flag_at || c.live_flag_clean < (size_t)c.S) {
    CAC_HIP(hipMemsetAsync(c.live_flag.get(), 0, c.S, c.stream));
    c.live_flag_at = c.live_flag.get();
  }
  c.live_flag_clean = 0;  // until this prune's flags are compacted (and so cleared)
  const int blocks = (int)std::min<long>(grid_for(16L * c.P, 256), 8L * c.ncu);
  const unsigned* nl = reinterpret_cast<const unsigned*>(c.nlist.get());
  unsigned* il = reinterpret_cast<unsigned*>(c.ilist.get());
  if (mixed)
    k_prune<true><<<blocks, 256, 0, c.stream>>>(c.site.get(), c.site32f.get(), c.row_site.get(), c.row_point.get(),
                                                 c.nnum.get(), nl, c.P, c.cap, g, c.inner_cutsq, c.innum.get(), il,
                                                 c.icap, c.x.get(), c.xhold_in.get(), c.live_flag.get());
  else
    k_prune<false><<<blocks, 256, 0, c.stream>>>(c.site.get(), c.site32f.get(), c.row_site.get(), c.row_point.get(),
                                                  c.nnum.get(), nl, c.P, c.cap, g, c.inner_cutsq, c.innum.get(), il,
                                                  c.icap, c.x.get(), c.xhold_in.get(), c.live_flag.get());
  compact_live(c);
  // the dummy entry: no cell offset (mixed) / no image shift (FP64)
  const unsigned pad_code = mixed ? (62u << MIX_SHIFT) : (13u << IMG_SHIFT);
  k_live_remap<<<blocks, 256, 0, c.stream>>>(c.live_idx.get(), c.row_site.get(), c.row_live.get(), c.innum.get(), il,
                                             c.P, c.icap, mixed ? (1u << MIX_SHIFT) - 1u : SITE_MASK, pad_code,
                                             c.d_nlive.get(), mixed ? c.site32.get() : nullptr,
                                             mixed ? nullptr : c.site_live.get());
  c.inner_mode = c.precision;
  c.inner_builds++;
}

// ------------------------------------------------------------------ export for parity checks

void export_lists(Ctx& c, std::vector<int>& row, std::vector<int>& kind, std::vector<int>& owner,
                  std::vector<int>& ucell, std::vector<int>& img) {
  const int P = c.P, S = c.S;
  std::vector<int> nnum(P), rp(P), rs(P), sdesc(S), simg(S);
  std::vector<int> nl((size_t)c.cap * P);
  CAC_HIP(hipStreamSynchronize(c.stream));
  CAC_HIP(hipMemcpy(nnum.data(), c.nnum.get(), sizeof(int) * P, hipMemcpyDeviceToHost));
  CAC_HIP(hipMemcpy(rp.data(), c.row_point.get(), sizeof(int) * P, hipMemcpyDeviceToHost));
  CAC_HIP(hipMemcpy(rs.data(), c.row_site.get(), sizeof(int) * P, hipMemcpyDeviceToHost));
  CAC_HIP(hipMemcpy(sdesc.data(), c.sdesc.get(), sizeof(int) * S, hipMemcpyDeviceToHost));
  CAC_HIP(hipMemcpy(simg.data(), c.simg.get(), sizeof(int) * S, hipMemcpyDeviceToHost));
  CAC_HIP(hipMemcpy(nl.data(), c.nlist.get(), sizeof(int) * (size_t)c.cap * P, hipMemcpyDeviceToHost));
  auto decode = [](int code, int* s) {
    s[0] = code % 3 - 1;
    s[1] = (code / 3) % 3 - 1;
    s[2] = code / 9 - 1;
  };
  row.clear(); kind.clear(); owner.clear(); ucell.clear(); img.clear();
  for (int r = 0; r < P; r++) {
    const int p = rp[r];
    int prow;
    if (p < c.A)
      prow = p;
    else {
      int e = (p - c.A) / NPE, k = (p - c.A) % NPE;
      prow = c.A + NPE * e + c.n2g[k];
    }
    int sp[3];
    decode(simg[rs[r]], sp);
    for (int n = 0; n < nnum[r]; n++) {
      const unsigned ent = (unsigned)nl[(size_t)r * c.cap + n];
      const int j = (int)(ent & SITE_MASK), t = (int)(ent >> IMG_SHIFT);
      int ts[3], ss[3];
      decode(t, ts);
      decode(simg[j], ss);
      const int d = sdesc[j];
      int kd, ow, uc;
      if (d >= 0 && d < c.A) {
        kd = 0; ow = d; uc = 0;
      } else if (d >= 0) {
        int e = (d - c.A) / NPE, k = (d - c.A) % NPE;
        kd = 1; ow = e; uc = c.g2u[c.n2g[k]];
      } else {
        int q = -d - 1;
        kd = 1; ow = q / c.nucell; uc = q % c.nucell;
      }
      row.push_back(prow);
      kind.push_back(kd);
      owner.push_back(ow);
      ucell.push_back(uc);
      for (int dd = 0; dd < 3; dd++) img.push_back(ss[dd] + ts[dd] - sp[dd]);
    }
  }
}

}  // namespace cac
