// pair.hip -- site refresh and the EAM (three passes) and LJ force kernels.
//
// Replaces PairEAM::compute (pair_eam.cpp:618-1032), PairEAM::compute_fp
// (pair_eam.cpp:1160-1169), PairLJCut::compute (pair_lj_cut.cpp:227-417) and the
// forward/reverse comm they rely on (CommBrick::forward_comm / reverse_comm,
// forward_comm_pair / reverse_comm_pair).
//
// Every point (atom or Gauss point) owns a row of the full neighbor list and
// sums its own density, force, energy and virial from that row: no scatter, no
// float atomics, fixed summation order, so results are bit-reproducible.  The
// reference's half list stores a pair once and applies the reaction to j only
// when j is an atom or a Gauss point; the full list yields the same forces,
// because every site that would receive a reaction is itself a row, and sites
// that would not (non-Gauss interpolated atoms) own no row.  Energies and
// virials carry the reference's per-side scale (1/2 for an atom, nodal_weight/2
// for a node, 0 otherwise; pair_eam.cpp:867-918, 979-984).
//
// Mixed precision: distances come from FP64 positions and are rounded to FP32;
// r, the spline evaluations and the pair term run in FP32; density, force,
// energy and virial accumulate in FP64.
#include "cac_ctx.hpp"

namespace cac {

struct PairArgs {
  const double* sx;    // [3][S_cap]
  const int* row_site;
  const int* row_point;
  const int* nnum;
  const int* nlist;
  int R, S_cap, A;
  double L[3];
  double cutsq;
};

__device__ __forceinline__ void img_shift(int code, const double* L, double& sx, double& sy, double& sz) {
  sx = (double)(code % 3 - 1) * L[0];
  sy = (double)((code / 3) % 3 - 1) * L[1];
  sz = (double)(code / 9 - 1) * L[2];
}

// ------------------------------------------------------------------ site refresh
// forward_comm equivalent: every site position from its source point or from
// the shape-function interpolation of its element's nodes, plus its wrap shift.
__global__ void k_site_refresh(const double* __restrict__ x, const double* __restrict__ shape,
                               const int* __restrict__ sdesc, const int* __restrict__ simg, int S, int S_cap, int A,
                               int P, int nucell, double L0, double L1, double L2, double* __restrict__ sx) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S) return;
  const int d = sdesc[k];
  const int code = simg[k];
  const double L[3] = {L0, L1, L2};
  double sh[3];
  img_shift(code, L, sh[0], sh[1], sh[2]);
  if (d >= 0) {
    for (int c = 0; c < 3; c++) sx[(size_t)c * S_cap + k] = x[(size_t)c * P + d] + sh[c];
  } else {
    const int q = -d - 1;
    const int e = q / nucell, u = q - e * nucell;
    const double* N = shape + (size_t)u * NPE;
    const size_t base = (size_t)A + (size_t)NPE * e;
    for (int c = 0; c < 3; c++) {
      const double* nx = x + (size_t)c * P + base;
      double s = 0.0;
      for (int n = 0; n < NPE; n++) s += N[n] * nx[n];
      sx[(size_t)c * S_cap + k] = s + sh[c];
    }
  }
}

void site_refresh(Ctx& c) {
  ScopedTimer tm(c, "site_refresh");
  k_site_refresh<<<grid_for(c.S, 256), 256, 0, c.stream>>>(c.x.get(), c.shape.get(), c.sdesc.get(), c.simg.get(),
                                                           c.S, c.S, c.A, c.P, c.nucell, c.L[0], c.L[1], c.L[2],
                                                           c.sx.get());
}

// ------------------------------------------------------------------ spline helpers
// knot index and fraction exactly as pair_eam.cpp:799-805 / 1160-1169

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

__device__ __forceinline__ int knot_rho(double rho, double rdrho, int nrho, double& p) {
  p = rho * rdrho + 1.0;
  int m = (int)p;
  m = m < nrho - 1 ? m : nrho - 1;
  m = m > 1 ? m : 1;
  p -= m;
  p = p < 1.0 ? p : 1.0;
  return m;
}

// ------------------------------------------------------------------ EAM density (pair_eam.cpp:721-813)

template <bool MIXED>
__global__ __launch_bounds__(256) void k_eam_density(PairArgs a, const double* __restrict__ rhor,
                                                     const float* __restrict__ rhor32, double rdr, int nr,
                                                     double* __restrict__ rho_out) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R) return;
  const int k = a.row_site[r];
  const double px = a.sx[k], py = a.sx[(size_t)a.S_cap + k], pz = a.sx[2 * (size_t)a.S_cap + k];
  const int n = a.nnum[r];
  double rho = 0.0;
  for (int jj = 0; jj < n; jj++) {
    const unsigned ent = (unsigned)a.nlist[(size_t)jj * a.R + r];
    const int j = (int)(ent & SITE_MASK);
    double shx, shy, shz;
    img_shift((int)(ent >> IMG_SHIFT), a.L, shx, shy, shz);
    const double delx = px - (a.sx[j] + shx);
    const double dely = py - (a.sx[(size_t)a.S_cap + j] + shy);
    const double delz = pz - (a.sx[2 * (size_t)a.S_cap + j] + shz);
    if (MIXED) {
      const float fx = (float)delx, fy = (float)dely, fz = (float)delz;
      const float rsq = fx * fx + fy * fy + fz * fz;
      if (rsq < (float)a.cutsq) {
        float p;
        const int m = knot_r<float>(sqrtf(rsq), (float)rdr, nr, p);
        const float* c = rhor32 + (size_t)m * 4;  // value coefficients 3..6
        rho += (double)(((c[0] * p + c[1]) * p + c[2]) * p + c[3]);
      }
    } else {
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < a.cutsq) {
        double p;
        const int m = knot_r<double>(sqrt(rsq), rdr, nr, p);
        const double* c = rhor + (size_t)m * 7;
        rho += ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
      }
    }
  }
  rho_out[a.row_point[r]] = rho;
}

// ------------------------------------------------------------------ embedding derivative per site
// atoms and nodes: F'(rho) of their own density (pair_eam.cpp:822-832, 910);
// interpolated atoms: F'(sum_k N_k rho_node_k), the value the reference
// recomputes for every pair (pair_eam.cpp:959-969)

__global__ void k_eam_embed(const double* __restrict__ rho, const double* __restrict__ shape,
                            const int* __restrict__ sdesc, int S, int A, int nucell,
                            const double* __restrict__ frho, double rdrho, int nrho, double* __restrict__ sfp,
                            float* __restrict__ sfp32) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S) return;
  const int d = sdesc[k];
  double rj;
  if (d >= 0) {
    rj = rho[d];
  } else {
    const int q = -d - 1;
    const int e = q / nucell, u = q - e * nucell;
    const double* N = shape + (size_t)u * NPE;
    const double* nr = rho + (size_t)A + (size_t)NPE * e;
    rj = 0.0;
    for (int n = 0; n < NPE; n++) rj += N[n] * nr[n];
  }
  double p;
  const int m = knot_rho(rj, rdrho, nrho, p);
  const double* c = frho + (size_t)m * 7;
  const double fp = (c[0] * p + c[1]) * p + c[2];
  sfp[k] = fp;
  sfp32[k] = (float)fp;
}

// ------------------------------------------------------------------ EAM force (pair_eam.cpp:853-1027)

struct EamForceArgs {
  const double* rhor;  // [(nr+1)*7]
  const double* z2r;
  const double* frho;
  const float* fr32;   // [(nr+1)*7]: rhor' coeffs 0..2, z2r coeffs 0..6 ... packed per knot
  const double* rho;   // per point
  const double* sfp;
  const float* sfp32;
  double rdr, rdrho, rhomax;
  int nr, nrho;
  double escale_atom, escale_node, vscale_node;
};

template <bool MIXED, bool EV>
__global__ __launch_bounds__(256) void k_eam_force(PairArgs a, EamForceArgs q, double* __restrict__ f,
                                                   double* __restrict__ gf, int P, int E8,
                                                   double* __restrict__ erow) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R) return;
  const int k = a.row_site[r];
  const int pnt = a.row_point[r];
  const bool is_atom = pnt < a.A;
  const double px = a.sx[k], py = a.sx[(size_t)a.S_cap + k], pz = a.sx[2 * (size_t)a.S_cap + k];
  const double fpi = q.sfp[k];
  const float fpi32 = q.sfp32[k];
  const int n = a.nnum[r];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  double ephi = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
  for (int jj = 0; jj < n; jj++) {
    const unsigned ent = (unsigned)a.nlist[(size_t)jj * a.R + r];
    const int j = (int)(ent & SITE_MASK);
    double shx, shy, shz;
    img_shift((int)(ent >> IMG_SHIFT), a.L, shx, shy, shz);
    const double delx = px - (a.sx[j] + shx);
    const double dely = py - (a.sx[(size_t)a.S_cap + j] + shy);
    const double delz = pz - (a.sx[2 * (size_t)a.S_cap + j] + shz);
    if (MIXED) {
      const float dx = (float)delx, dy = (float)dely, dz = (float)delz;
      const float rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < (float)a.cutsq) {
        const float rr = sqrtf(rsq);
        float p;
        const int m = knot_r<float>(rr, (float)q.rdr, q.nr, p);
        const float* c = q.fr32 + (size_t)m * 10;
        const float rhop = (c[0] * p + c[1]) * p + c[2];
        const float z2p = (c[3] * p + c[4]) * p + c[5];
        const float z2 = ((c[6] * p + c[7]) * p + c[8]) * p + c[9];
        const float recip = 1.0f / rr;
        const float phi = z2 * recip;
        const float phip = z2p * recip - phi * recip;
        const float psip = fpi32 * rhop + q.sfp32[j] * rhop + phip;
        const float fpair = -psip * recip;
        fx += (double)(dx * fpair);
        fy += (double)(dy * fpair);
        fz += (double)(dz * fpair);
        if (EV) {
          ephi += (double)phi;
          v0 += (double)(dx * dx * fpair);
          v1 += (double)(dy * dy * fpair);
          v2 += (double)(dz * dz * fpair);
          v3 += (double)(dx * dy * fpair);
          v4 += (double)(dx * dz * fpair);
          v5 += (double)(dy * dz * fpair);
        }
      }
    } else {
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < a.cutsq) {
        const double rr = sqrt(rsq);
        double p;
        const int m = knot_r<double>(rr, q.rdr, q.nr, p);
        const double* c = q.rhor + (size_t)m * 7;
        const double rhop = (c[0] * p + c[1]) * p + c[2];
        const double* z = q.z2r + (size_t)m * 7;
        const double z2p = (z[0] * p + z[1]) * p + z[2];
        const double z2 = ((z[3] * p + z[4]) * p + z[5]) * p + z[6];
        const double recip = 1.0 / rr;
        const double phi = z2 * recip;
        const double phip = z2p * recip - phi * recip;
        const double psip = fpi * rhop + q.sfp[j] * rhop + phip;
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
      }
    }
  }
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
    double p;
    const int m = knot_rho(rho, q.rdrho, q.nrho, p);
    const double* c = q.frho + (size_t)m * 7;
    double phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
    if (rho > q.rhomax) phi += fpi * (rho - q.rhomax);
    const double es = is_atom ? q.escale_atom : q.escale_node;
    const double vs = is_atom ? 0.5 : q.vscale_node;
    erow[r] = phi * (is_atom ? 1.0 : 2.0 * q.escale_node) + es * ephi;
    erow[(size_t)a.R * 1 + r] = vs * v0;
    erow[(size_t)a.R * 2 + r] = vs * v1;
    erow[(size_t)a.R * 3 + r] = vs * v2;
    erow[(size_t)a.R * 4 + r] = vs * v3;
    erow[(size_t)a.R * 5 + r] = vs * v4;
    erow[(size_t)a.R * 6 + r] = vs * v5;
  }
}

// ------------------------------------------------------------------ LJ (pair_lj_cut.cpp:278-414)

struct LjArgs {
  double lj1, lj2, lj3, lj4, escale_node, vscale_node;
};

template <bool MIXED, bool EV>
__global__ __launch_bounds__(256) void k_lj_force(PairArgs a, LjArgs q, double* __restrict__ f,
                                                  double* __restrict__ gf, int P, int E8,
                                                  double* __restrict__ erow) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.R) return;
  const int k = a.row_site[r];
  const int pnt = a.row_point[r];
  const bool is_atom = pnt < a.A;
  const double px = a.sx[k], py = a.sx[(size_t)a.S_cap + k], pz = a.sx[2 * (size_t)a.S_cap + k];
  const int n = a.nnum[r];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  double ev = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
  const float l1 = (float)q.lj1, l2 = (float)q.lj2, l3 = (float)q.lj3, l4 = (float)q.lj4;
  for (int jj = 0; jj < n; jj++) {
    const unsigned ent = (unsigned)a.nlist[(size_t)jj * a.R + r];
    const int j = (int)(ent & SITE_MASK);
    double shx, shy, shz;
    img_shift((int)(ent >> IMG_SHIFT), a.L, shx, shy, shz);
    const double delx = px - (a.sx[j] + shx);
    const double dely = py - (a.sx[(size_t)a.S_cap + j] + shy);
    const double delz = pz - (a.sx[2 * (size_t)a.S_cap + j] + shz);
    if (MIXED) {
      const float dx = (float)delx, dy = (float)dely, dz = (float)delz;
      const float rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < (float)a.cutsq) {
        const float r2inv = 1.0f / rsq;
        const float r6inv = r2inv * r2inv * r2inv;
        const float forcelj = r6inv * (l1 * r6inv - l2);
        const float fpair = forcelj * r2inv;
        fx += (double)(dx * fpair);
        fy += (double)(dy * fpair);
        fz += (double)(dz * fpair);
        if (EV) {
          ev += (double)(r6inv * (l3 * r6inv - l4));
          v0 += (double)(dx * dx * fpair);
          v1 += (double)(dy * dy * fpair);
          v2 += (double)(dz * dz * fpair);
          v3 += (double)(dx * dy * fpair);
          v4 += (double)(dx * dz * fpair);
          v5 += (double)(dy * dz * fpair);
        }
      }
    } else {
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < a.cutsq) {
        const double r2inv = 1.0 / rsq;
        const double r6inv = r2inv * r2inv * r2inv;
        const double forcelj = r6inv * (q.lj1 * r6inv - q.lj2);
        const double fpair = forcelj * r2inv;
        fx += delx * fpair;
        fy += dely * fpair;
        fz += delz * fpair;
        if (EV) {
          ev += r6inv * (q.lj3 * r6inv - q.lj4);
          v0 += delx * delx * fpair;
          v1 += dely * dely * fpair;
          v2 += delz * delz * fpair;
          v3 += delx * dely * fpair;
          v4 += delx * delz * fpair;
          v5 += dely * delz * fpair;
        }
      }
    }
  }
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
  }
}

// ------------------------------------------------------------------ orchestration

static PairArgs pair_args(Ctx& c, double cutsq) {
  PairArgs a;
  a.sx = c.sx.get();
  a.row_site = c.row_site.get();
  a.row_point = c.row_point.get();
  a.nnum = c.nnum.get();
  a.nlist = c.nlist.get();
  a.R = c.P;
  a.S_cap = c.S;
  a.A = c.A;
  for (int d = 0; d < 3; d++) a.L[d] = c.L[d];
  a.cutsq = cutsq;
  return a;
}

void pair_compute(Ctx& c, int eflag, int vflag) {
  const int R = c.P;
  const bool ev = eflag || vflag;
  const bool mixed = c.precision == 1;
  if (ev) c.erow.reserve((size_t)7 * R);
  const int grid = grid_for(R, 256);
  const int E8 = NPE * c.E;
  double* gf = c.gf.get();
  if (c.pairstyle == 2) {
    PairArgs a = pair_args(c, c.cutforcesq);
    {
      ScopedTimer tm(c, "eam_density");
      if (mixed)
        k_eam_density<true><<<grid, 256, 0, c.stream>>>(a, c.rhor_spl.get(), c.rhor_spl32.get(), c.rdr, c.nr,
                                                        c.rho.get());
      else
        k_eam_density<false><<<grid, 256, 0, c.stream>>>(a, c.rhor_spl.get(), c.rhor_spl32.get(), c.rdr, c.nr,
                                                         c.rho.get());
    }
    {
      ScopedTimer tm(c, "eam_embed");
      k_eam_embed<<<grid_for(c.S, 256), 256, 0, c.stream>>>(c.rho.get(), c.shape.get(), c.sdesc.get(), c.S, c.A,
                                                            c.nucell > 0 ? c.nucell : 1, c.frho_spl.get(), c.rdrho,
                                                            c.nrho, c.sfp.get(), c.sfp32.get());
    }
    EamForceArgs q;
    q.rhor = c.rhor_spl.get();
    q.z2r = c.z2r_spl.get();
    q.frho = c.frho_spl.get();
    q.fr32 = c.z2r_spl32.get();
    q.rho = c.rho.get();
    q.sfp = c.sfp.get();
    q.sfp32 = c.sfp32.get();
    q.rdr = c.rdr;
    q.rdrho = c.rdrho;
    q.rhomax = c.rhomax;
    q.nr = c.nr;
    q.nrho = c.nrho;
    q.escale_atom = 0.5;
    q.escale_node = c.nodal_weight / 2.0;
    q.vscale_node = c.nodal_weight / 2.0;
    ScopedTimer tm(c, "eam_force");
    if (mixed) {
      if (ev)
        k_eam_force<true, true><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, c.erow.get());
      else
        k_eam_force<true, false><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, nullptr);
    } else {
      if (ev)
        k_eam_force<false, true><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, c.erow.get());
      else
        k_eam_force<false, false><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, nullptr);
    }
  } else {
    PairArgs a = pair_args(c, c.ljcutsq);
    LjArgs q{c.lj1, c.lj2, c.lj3, c.lj4, c.nodal_weight / 2.0, c.nodal_weight / 2.0};
    ScopedTimer tm(c, "lj_force");
    if (mixed) {
      if (ev)
        k_lj_force<true, true><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, c.erow.get());
      else
        k_lj_force<true, false><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, nullptr);
    } else {
      if (ev)
        k_lj_force<false, true><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, c.erow.get());
      else
        k_lj_force<false, false><<<grid, 256, 0, c.stream>>>(a, q, c.f.get(), gf, c.P, E8, nullptr);
    }
  }
  if (c.E) nodal_assembly(c);
  if (ev) {
    double out[7];
    reduce_rows(c, c.erow.get(), R, 7, out);
    if (eflag) c.eng_vdwl = out[0];
    if (vflag)
      for (int i = 0; i < 6; i++) c.virial[i] = out[1 + i];
  }
}

}  // namespace cac
