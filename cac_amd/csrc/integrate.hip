// integrate.hip -- velocity Verlet, element centres, nodal force assembly and
// deterministic reductions.
//
// Replaces FixNVE::initial_integrate / final_integrate (fix_nve.cpp:62-217),
// ElementVecCAC::update_center_coord (element_vec_cac.cpp:3740-3765),
// Element::compute_nodef (element.cpp:2178-2246) and the thermo sums of
// ComputeTemp / ComputePE / ComputePressure (compute_temp.cpp:73-157).
#include <map>
#include <mutex>

#include "cac_ctx.hpp"

namespace cac {

struct MassTab {
  double inv[MAXT + 1];  // dtf / mass[type] is formed per point as the reference does
  double m[MAXT + 1];
};

// atoms and nodes move the same way; nodes use mass[ctype] (fix_nve.cpp:118-131)
__global__ void k_nve_initial(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ f,
                              const int* __restrict__ ptype, int P, MassTab mt, double dtf, double dtv) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double dtfm = dtf / mt.m[ptype[p]];
  for (int d = 0; d < 3; d++) {
    const size_t i = (size_t)d * P + p;
    const double vv = v[i] + dtfm * f[i];
    v[i] = vv;
    x[i] = x[i] + dtv * vv;
  }
}

__global__ void k_nve_final(double* __restrict__ v, const double* __restrict__ f, const int* __restrict__ ptype,
                            int P, MassTab mt, double dtf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double dtfm = dtf / mt.m[ptype[p]];
  for (int d = 0; d < 3; d++) {
    const size_t i = (size_t)d * P + p;
    v[i] = v[i] + dtfm * f[i];
  }
}

// element centre = mean of its nodes, summed in node order then / npe / apc
__global__ void k_centers(const double* __restrict__ x, double* __restrict__ ex, int A, int E, int P) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  for (int d = 0; d < 3; d++) {
    const double* nx = x + (size_t)d * P + A + (size_t)NPE * e;
    double s = 0.0;
    for (int k = 0; k < NPE; k++) s += nx[k];
    ex[(size_t)d * E + e] = s / NPE / 1;
  }
}

static MassTab mass_tab(const Ctx& c) {
  MassTab t;
  for (int i = 0; i <= MAXT; i++) {
    t.m[i] = c.mass[i] > 0 ? c.mass[i] : 1.0;
    t.inv[i] = 1.0 / t.m[i];
  }
  return t;
}

// element centres are needed only where Domain::pbc wraps elements, at a
// neighbor build, which recomputes them from the nodes there
void nve_initial(Ctx& c, double dt) {
  ScopedTimer tm(c, "nve");
  const double dtf = 0.5 * dt * c.ftm2v;
  k_nve_initial<<<grid_for(c.P, 256), 256, 0, c.stream>>>(c.x.get(), c.v.get(), c.f.get(), c.ptype.get(), c.P,
                                                          mass_tab(c), dtf, dt);
  c.x_dirty = true;
}

void nve_final(Ctx& c, double dt) {
  ScopedTimer tm(c, "nve");
  const double dtf = 0.5 * dt * c.ftm2v;
  k_nve_final<<<grid_for(c.P, 256), 256, 0, c.stream>>>(c.v.get(), c.f.get(), c.ptype.get(), c.P, mass_tab(c), dtf);
}

void update_centers(Ctx& c) {
  if (!c.E) return;
  k_centers<<<grid_for(c.E, 256), 256, 0, c.stream>>>(c.x.get(), c.ex.get(), c.A, c.E, c.P);
}

// ------------------------------------------------------------------ nodal assembly
// force_columns[k] = sum_g wsa[g][k] gaussf[g]; nodef[a] = sum_b Minv[a][b] fc[b]
// (CONSISTENT), nodef = fc (LUMPED, with the lumped weights) or the Gauss-cell
// force itself (NODEONLY).  One thread per element and dimension; no atomics.

struct NodefTab {
  double w[8][8];     // [g][node]
  double minv[8][8];  // [a][b]
  int g2n[8], n2g[8];
  int style;
};

__global__ void k_nodef(const double* __restrict__ gf, double* __restrict__ f, int A, int E, int P, NodefTab t) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 3 * E) return;
  const int d = id / E, e = id - d * E;
  const int E8 = NPE * E;
  // Gauss-cell forces of this element, indexed by Gauss cell g (stored at node g2n[g])
  double g[8];
  for (int q = 0; q < 8; q++) g[q] = gf[(size_t)d * E8 + (size_t)NPE * e + t.g2n[q]];
  double* out = f + (size_t)d * P + A + (size_t)NPE * e;
  if (t.style == 2) {
    for (int k = 0; k < NPE; k++) out[k] = g[t.n2g[k]];
    return;
  }
  double fc[8];
  for (int k = 0; k < NPE; k++) {
    double s = 0.0;
    for (int q = 0; q < 8; q++) s += t.w[q][k] * g[q];
    fc[k] = s;
  }
  if (t.style == 1) {
    for (int k = 0; k < NPE; k++) out[k] = fc[k];
    return;
  }
  for (int a = 0; a < NPE; a++) {
    double s = 0.0;
    for (int b = 0; b < NPE; b++) s += t.minv[a][b] * fc[b];
    out[a] = s;
  }
}

static NodefTab nodef_tab(const Ctx& c) {
  NodefTab t;
  for (int g = 0; g < 8; g++)
    for (int k = 0; k < 8; k++) {
      t.w[g][k] = c.nodef_style == 1 ? c.wsa_lumped[g][k] : c.wsa[g][k];
      t.minv[g][k] = c.minv[g][k];
    }
  for (int k = 0; k < 8; k++) {
    t.g2n[k] = c.g2n[k];
    t.n2g[k] = c.n2g[k];
  }
  t.style = c.nodef_style;
  return t;
}

void nodal_assembly(Ctx& c) {
  ScopedTimer tm(c, "nodef");
  k_nodef<<<grid_for(3L * c.E, 256), 256, 0, c.stream>>>(c.gf.get(), c.f.get(), c.A, c.E, c.P, nodef_tab(c));
}

// ------------------------------------------------------------------ fused step ends
// Kick + drift of every point (fix_nve.cpp:62-131) and, on the new positions,
// the reference's trigger (any point moved more than skin/2 since the last
// build, neighbor.cpp:1617-1683) and the inner list's (moved more than
// delta/2 since the last prune).  A wave with a hit writes the step number
// into the device and host-mapped step words; every writer writes the same
// value, so plain stores suffice.
// The check's verdict is complete once every workgroup of the launch has
// written its step words.  A wave that writes a step word fences at system
// scope right after (only on a step whose check fires).  Each workgroup then
// counts itself on a device counter after a barrier; the workgroup that
// completes the launch's count (the host passes the running total it
// expects) fences and writes the step number into the host-mapped `done`
// word, which the host polls instead of waiting on an event recorded in the
// stream.
// The count has two levels so that few workgroups update any one address:
// workgroup b counts itself on the counter of XCD b % 8, and the workgroup
// that completes an XCD's count then counts that XCD on a global counter.
// Counters sit 256 bytes apart; the host passes the running targets.
struct CheckTargets {
  unsigned xcd[8];  // workgroups counted on each XCD counter so far
  unsigned all;     // XCD counts completed so far
};
constexpr int CHECK_STRIDE = 64;  // unsigned words between counters

__device__ __forceinline__ void publish_check(unsigned* count, const CheckTargets& t, int* done, int step) {
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int x = blockIdx.x % 8;
  if (atomicAdd(count + CHECK_STRIDE * x, 1u) != t.xcd[x] - 1u) return;
  if (atomicAdd(count + CHECK_STRIDE * 8, 1u) != t.all - 1u) return;
  __threadfence_system();
  *reinterpret_cast<volatile int*>(done) = step;
}

// the running targets after one more launch of `blocks` workgroups
static CheckTargets next_targets(Ctx& c, int blocks) {
  for (int x = 0; x < 8; x++) {
    const unsigned nx = (unsigned)((blocks - x + 7) / 8);  // workgroups b < blocks with b % 8 == x
    c.check_xcd[x] += nx;
    if (nx) c.check_all++;
  }
  CheckTargets t;
  for (int x = 0; x < 8; x++) t.xcd[x] = c.check_xcd[x];
  t.all = c.check_all;
  return t;
}

// the step words of a wave whose points hit a trigger
__device__ __forceinline__ void mark_step(int* d_word, int* h_word, int step) {
  *d_word = step;
  *h_word = step;
  __threadfence_system();
}

// workgroup size of the fused kick / check kernels (cac_eam_x8,
// k_step_end_begin: 16.4 us with the two-level count; with a single counter
// 27.0 us at 256 threads, 18.9 at 512 and 20.4 at 1024; 16.6 at 256 without
// any count)
constexpr int CB = 512;

__global__ __launch_bounds__(CB) void k_step_begin(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ f,
                             const int* __restrict__ ptype, const double* __restrict__ xh,
                             const double* __restrict__ xhi, int P, MassTab mt, double dtf, double dtv, double trig,
                             double trig_in, int outer, int step, int* __restrict__ d_steps,
                             int* __restrict__ h_steps, unsigned* __restrict__ count, CheckTargets target) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int hit = 0, hit_in = 0;
  if (p < P) {
    const double dtfm = dtf / mt.m[ptype[p]];
    double xn[3];
    for (int d = 0; d < 3; d++) {
      const size_t i = (size_t)d * P + p;
      const double vv = v[i] + dtfm * f[i];
      v[i] = vv;
      xn[d] = x[i] + dtv * vv;
      x[i] = xn[d];
    }
    if (outer) {
      const double dx = xn[0] - xh[p], dy = xn[1] - xh[(size_t)P + p], dz = xn[2] - xh[2 * (size_t)P + p];
      hit = (dx * dx + dy * dy + dz * dz > trig);
    }
    const double ex = xn[0] - xhi[p], ey = xn[1] - xhi[(size_t)P + p], ez = xn[2] - xhi[2 * (size_t)P + p];
    hit_in = (ex * ex + ey * ey + ez * ez > trig_in);
  }
  const bool lead = (threadIdx.x & 63) == 0;
  if (__ballot(hit) && lead) mark_step(d_steps, h_steps, step);
  if (__ballot(hit_in) && lead) mark_step(d_steps + 1, h_steps + 1, step);
  publish_check(count, target, h_steps + 2, step);
}

void step_begin(Ctx& c, double dt, int step, bool outer) {
  ScopedTimer tm(c, "nve");
  const double dtf = 0.5 * dt * c.ftm2v;
  const double hd = 0.5 * c.inner_skin;
  const int blocks = grid_for(c.P, CB);
  const CheckTargets t = next_targets(c, blocks);
  k_step_begin<<<blocks, CB, 0, c.stream>>>(c.x.get(), c.v.get(), c.f.get(), c.ptype.get(), c.xhold.get(),
                                             c.xhold_in.get(), c.P, mass_tab(c), dtf, dt, c.triggersq, hd * hd,
                                             outer ? 1 : 0, step, c.d_steps.get(), c.h_steps_dev,
                                             c.d_check_count.get(), t);
  c.x_dirty = true;
}

// Nodal forces of every element (Element::compute_nodef) and the final kick of
// every point (fix_nve.cpp:139-217): threads [0, A) kick atoms; thread A + d E + e
// assembles dimension d of element e's nodal forces and kicks its nodes.
__global__ void k_step_end(const double* __restrict__ gf, double* __restrict__ f, double* __restrict__ v,
                           const int* __restrict__ ptype, int A, int E, int P, NodefTab t, MassTab mt, double dtf,
                           Gate g) {
  if (g.skip()) return;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < A) {
    const double dtfm = dtf / mt.m[ptype[id]];
    for (int d = 0; d < 3; d++) {
      const size_t i = (size_t)d * P + id;
      v[i] = v[i] + dtfm * f[i];
    }
    return;
  }
  const int q = id - A;
  if (q >= 3 * E) return;
  const int d = q / E, e = q - d * E;
  const int E8 = NPE * E;
  double gq[8];
  for (int k = 0; k < 8; k++) gq[k] = gf[(size_t)d * E8 + (size_t)NPE * e + t.g2n[k]];
  double out[NPE];
  if (t.style == 2) {
    for (int k = 0; k < NPE; k++) out[k] = gq[t.n2g[k]];
  } else {
    double fc[NPE];
    for (int k = 0; k < NPE; k++) {
      double s = 0.0;
      for (int j = 0; j < 8; j++) s += t.w[j][k] * gq[j];
      fc[k] = s;
    }
    if (t.style == 1) {
      for (int k = 0; k < NPE; k++) out[k] = fc[k];
    } else {
      for (int a = 0; a < NPE; a++) {
        double s = 0.0;
        for (int b = 0; b < NPE; b++) s += t.minv[a][b] * fc[b];
        out[a] = s;
      }
    }
  }
  const size_t base = (size_t)d * P + A + (size_t)NPE * e;
  for (int k = 0; k < NPE; k++) {
    f[base + k] = out[k];
    const double dtfm = dtf / mt.m[ptype[A + (size_t)NPE * e + k]];
    v[base + k] = v[base + k] + dtfm * out[k];
  }
}

void step_end(Ctx& c, double dt, Gate g) {
  ScopedTimer tm(c, "nodef");
  const double dtf = 0.5 * dt * c.ftm2v;
  const long n = (long)c.A + 3L * c.E;
  k_step_end<<<grid_for(n, 256), 256, 0, c.stream>>>(c.gf.get(), c.f.get(), c.v.get(), c.ptype.get(), c.A, c.E, c.P,
                                                     nodef_tab(c), mass_tab(c), dtf, g);
}

// The closing half of step N and the opening half of step N + 1 in one pass
// over the points, for a step whose full-step velocities nobody reads (no
// thermo output): nodal forces of every element, then per point the final
// kick of N, the initial kick and drift of N + 1 and both displacement
// checks of N + 1.  Each point's arithmetic is that of k_step_end followed by
// k_step_begin, in the same order, so the state is bitwise the same.  Threads
// [0, A) are atoms; thread A + 8e + k is node k of element e, which forms its
// own nodal force in all three dimensions from the element's Gauss forces.
__global__ __launch_bounds__(CB) void k_step_end_begin(const double* __restrict__ gf, double* __restrict__ f, double* __restrict__ x,
                                 double* __restrict__ v, const int* __restrict__ ptype, const double* __restrict__ xh,
                                 const double* __restrict__ xhi, int A, int E, int P, NodefTab t, MassTab mt,
                                 double dtf, double dtv, double trig, double trig_in, int outer, int step,
                                 int* __restrict__ d_steps, int* __restrict__ h_steps, unsigned* __restrict__ count,
                                 CheckTargets target) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  int hit = 0, hit_in = 0;
  if (p < P) {
    double fp[3];
    if (p < A) {
      for (int d = 0; d < 3; d++) fp[d] = f[(size_t)d * P + p];
    } else {
      const int e = (p - A) / NPE, k = (p - A) - NPE * e;
      const int E8 = NPE * E;
      for (int d = 0; d < 3; d++) {
        double gq[8];
        for (int j = 0; j < 8; j++) gq[j] = gf[(size_t)d * E8 + (size_t)NPE * e + t.g2n[j]];
        double out;
        if (t.style == 2) {
          out = gq[t.n2g[k]];
        } else {
          double fc[NPE];
          for (int b = 0; b < NPE; b++) {
            double sacc = 0.0;
            for (int j = 0; j < 8; j++) sacc += t.w[j][b] * gq[j];
            fc[b] = sacc;
          }
          if (t.style == 1) {
            out = fc[k];
          } else {
            double sacc = 0.0;
            for (int b = 0; b < NPE; b++) sacc += t.minv[k][b] * fc[b];
            out = sacc;
          }
        }
        f[(size_t)d * P + p] = out;
        fp[d] = out;
      }
    }
    const double dtfm = dtf / mt.m[ptype[p]];
    double xn[3];
    for (int d = 0; d < 3; d++) {
      const size_t i = (size_t)d * P + p;
      const double vh = v[i] + dtfm * fp[d];   // final kick of step N
      const double vv = vh + dtfm * fp[d];     // initial kick of step N + 1
      v[i] = vv;
      xn[d] = x[i] + dtv * vv;
      x[i] = xn[d];
    }
    if (outer) {
      const double dx = xn[0] - xh[p], dy = xn[1] - xh[(size_t)P + p], dz = xn[2] - xh[2 * (size_t)P + p];
      hit = (dx * dx + dy * dy + dz * dz > trig);
    }
    const double ex = xn[0] - xhi[p], ey = xn[1] - xhi[(size_t)P + p], ez = xn[2] - xhi[2 * (size_t)P + p];
    hit_in = (ex * ex + ey * ey + ez * ez > trig_in);
  }
  const bool lead = (threadIdx.x & 63) == 0;
  if (__ballot(hit) && lead) mark_step(d_steps, h_steps, step);
  if (__ballot(hit_in) && lead) mark_step(d_steps + 1, h_steps + 1, step);
  publish_check(count, target, h_steps + 2, step);
}

void step_end_begin(Ctx& c, double dt, int next_step, bool outer) {
  ScopedTimer tm(c, "nve");
  const double dtf = 0.5 * dt * c.ftm2v;
  const double hd = 0.5 * c.inner_skin;
  const int blocks = grid_for(c.P, CB);
  const CheckTargets t = next_targets(c, blocks);
  k_step_end_begin<<<blocks, CB, 0, c.stream>>>(
      c.gf.get(), c.f.get(), c.x.get(), c.v.get(), c.ptype.get(), c.xhold.get(), c.xhold_in.get(), c.A, c.E, c.P,
      nodef_tab(c), mass_tab(c), dtf, dt, c.triggersq, hd * hd, outer ? 1 : 0, next_step, c.d_steps.get(),
      c.h_steps_dev, c.d_check_count.get(), t);
  c.x_dirty = true;
  c.begun_step = next_step;
}

// ------------------------------------------------------------------ deterministic reductions
// Each block sums a fixed slice in a fixed tree; a single block then sums the
// block partials in order.  No atomics: the result is bitwise reproducible.

constexpr int RB = 256;

__global__ __launch_bounds__(RB) void k_reduce_partial(const double* __restrict__ data, int n, int ncomp,
                                                       double* __restrict__ partial) {
  __shared__ double s[RB];
  const int nb = gridDim.x;
  for (int c = 0; c < ncomp; c++) {
    double acc = 0.0;
    for (long i = (long)blockIdx.x * RB + threadIdx.x; i < n; i += (long)nb * RB) acc += data[(size_t)c * n + i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = RB / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)c * nb + blockIdx.x] = s[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(RB) void k_reduce_final(const double* __restrict__ partial, int nb, int ncomp,
                                                     double* __restrict__ out) {
  __shared__ double s[RB];
  for (int c = 0; c < ncomp; c++) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) acc += partial[(size_t)c * nb + i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = RB / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s[0];
    __syncthreads();
  }
}

void reduce_rows(Ctx& c, const double* data, int n, int ncomp, double* out) {
  const int nb = 256;
  c.partials.reserve((size_t)nb * ncomp + 16);
  double* fin = c.partials.get() + (size_t)nb * ncomp;
  k_reduce_partial<<<nb, RB, 0, c.stream>>>(data, n, ncomp, c.partials.get());
  k_reduce_final<<<1, RB, 0, c.stream>>>(c.partials.get(), nb, ncomp, fin);
  CAC_HIP(hipMemcpyAsync(c.h_red, fin, sizeof(double) * ncomp, hipMemcpyDeviceToHost, c.stream));
  CAC_HIP(hipStreamSynchronize(c.stream));
  for (int i = 0; i < ncomp; i++) out[i] = c.h_red[i];
}

// ------------------------------------------------------------------ kinetic energy (compute_temp.cpp:96-150)

struct KeTab {
  double m[MAXT + 1];
  double mm[8][8];
  double nucell, nodal_weight;
  int consistent;
};

// m v^2 of row i: atom i < A, or element i - A with its consistent mass
// matrix times m nucell (or its nodes times m nodal_weight without one)
__device__ __forceinline__ double ke_row(const double* __restrict__ v, const int* __restrict__ ptype, int A, int P,
                                         const KeTab& t, int i) {
  if (i < A) {
    const double vx = v[i], vy = v[(size_t)P + i], vz = v[2 * (size_t)P + i];
    return t.m[ptype[i]] * (vx * vx + vy * vy + vz * vz);
  }
  const int e = i - A;
  const size_t b = (size_t)A + (size_t)NPE * e;
  double sum = 0.0;
  if (t.consistent) {
    for (int k = 0; k < NPE; k++)
      for (int l = 0; l < NPE; l++)
        sum += t.mm[k][l] * (v[b + k] * v[b + l] + v[(size_t)P + b + k] * v[(size_t)P + b + l] +
                             v[2 * (size_t)P + b + k] * v[2 * (size_t)P + b + l]);
    return sum * t.m[ptype[b]] * t.nucell;
  }
  for (int k = 0; k < NPE; k++)
    sum += v[b + k] * v[b + k] + v[(size_t)P + b + k] * v[(size_t)P + b + k] +
           v[2 * (size_t)P + b + k] * v[2 * (size_t)P + b + k];
  return sum * t.m[ptype[b]] * t.nodal_weight;
}

__global__ void k_ke_rows(const double* __restrict__ v, const int* __restrict__ ptype, int A, int E, int P, KeTab t,
                          double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A + E) return;
  out[i] = ke_row(v, ptype, A, P, t, i);
}

static KeTab ke_tab(const Ctx& c) {
  KeTab t;
  for (int i = 0; i <= MAXT; i++) t.m[i] = c.mass[i];
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++) t.mm[a][b] = c.mmat[a][b];
  t.nucell = c.nucell;
  t.nodal_weight = c.nodal_weight;
  t.consistent = c.nodef_style == 0;
  return t;
}

// temperature and kinetic energy from sum m v^2 (compute_temp.cpp:73-95, 140-157)
static void ke_temp(const Ctx& c, double s, double* ke, double* temp) {
  const double dof = 3.0 * ((double)c.A + (double)c.nucell * c.E);
  const double tfactor = dof > 0 ? c.mvv2e / (dof * c.boltz) : 0.0;
  const double T = s * tfactor;
  if (temp) *temp = T;
  if (ke) *ke = T * 0.5 * dof * c.boltz;
}

void thermo(Ctx& c, double* ke, double* temp) {
  const int n = c.A + c.E;
  c.erow.reserve((size_t)7 * c.P);
  k_ke_rows<<<grid_for(n, 256), 256, 0, c.stream>>>(c.v.get(), c.ptype.get(), c.A, c.E, c.P, ke_tab(c),
                                                   c.erow.get());
  double s = 0.0;
  reduce_rows(c, c.erow.get(), n, 1, &s);
  ke_temp(c, s, ke, temp);
}

// ------------------------------------------------------------------ thermo without device allocations
// The kinetic-energy sum of thermo() without the per-row buffer and the block
// partials the context grows on demand: each workgroup sums the m v^2 of a
// fixed slice of KE_RPB rows in a fixed tree and writes one partial into
// `scratch`, a file-scope device array; one workgroup then sums the partials
// in order into the slot after them, which is copied to pinned host memory
// the context allocated when it was created.  Every GPU holds its own copy of
// the array until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returned on that device (scratch_on).  Calls
// that overlap on one GPU share the array and race; each call writes every
// partial it reads before it reads it.
constexpr int KE_RPB = 256 * 16;           // rows per workgroup
constexpr int KE_MAX_BLOCKS = 4096;        // partials the array holds
constexpr long THERMO_NOALLOC_MAX_ROWS = (long)KE_RPB * KE_MAX_BLOCKS;  // 16,777,216 atoms + elements
static __device__ __align__(256) double scratch[KE_MAX_BLOCKS + 64];

__global__ __launch_bounds__(RB) void k_ke_partial(const double* __restrict__ v, const int* __restrict__ ptype, int A,
                                                   int E, int P, KeTab t, double* __restrict__ partial) {
  __shared__ double s[RB];
  const long first = (long)blockIdx.x * KE_RPB;
  const long last = first + KE_RPB < (long)A + E ? first + KE_RPB : (long)A + E;
  double acc = 0.0;
  for (long i = first + threadIdx.x; i < last; i += RB) acc += ke_row(v, ptype, A, P, t, (int)i);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// the device address of `scratch` on each device, taken once per device
static std::mutex scratch_mu;
static std::map<int, double*> scratch_addr;

double* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  auto it = scratch_addr.find(device);
  if (it != scratch_addr.end()) return it->second;
  int cur = 0;
  CAC_HIP(hipGetDevice(&cur));
  if (cur != device) throw HipError{hipErrorInvalidDevice, "scratch_on: the device must be current"};
  double* ptr = nullptr;
  CAC_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
  scratch_addr[device] = ptr;
  return ptr;
}

void thermo_noalloc(Ctx& c, double* ke, double* temp) {
  const long n = (long)c.A + c.E;
  if (n > THERMO_NOALLOC_MAX_ROWS)
    throw HipError{hipErrorInvalidValue, "thermo without allocation: more than 16,777,216 atoms + elements"};
  double* part = c.scratch;
  if (!part) throw HipError{hipErrorInvalidValue, "thermo without allocation: no scratch address on this device"};
  double s = 0.0;
  if (n > 0) {
    const int nb = (int)((n + KE_RPB - 1) / KE_RPB);
    k_ke_partial<<<nb, RB, 0, c.stream>>>(c.v.get(), c.ptype.get(), c.A, c.E, c.P, ke_tab(c), part);
    k_reduce_final<<<1, RB, 0, c.stream>>>(part, nb, 1, part + nb);
    CAC_HIP(hipGetLastError());
    CAC_HIP(hipMemcpyAsync(c.h_red, part + nb, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CAC_HIP(hipStreamSynchronize(c.stream));
    s = c.h_red[0];
  }
  ke_temp(c, s, ke, temp);
}

}  // namespace cac
