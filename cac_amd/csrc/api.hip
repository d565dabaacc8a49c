// api.hip -- the C ABI of include/cac_hip.h over the device state in cac_ctx.hpp.
//
// Host-side setup mirrors the reference's: element type tables
// (ElementVecCAC::set_element_types and friends), the EAM spline build
// (PairEAM::array2spline / interpolate), LJ coefficients (PairLJCut::init_one),
// neighbor cutoffs (Neighbor::init) and the Verlet step order (verlet.cpp:65-360).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/cac_hip.h"
#include "cac_ctx.hpp"
#include "eam_tables.hpp"

using namespace cac;

namespace cac {

Ctx::~Ctx() {
  if (h_red) (void)hipHostFree(h_red);
  if (h_flags) (void)hipHostFree(h_flags);
  if (h_steps) (void)hipHostFree(const_cast<int*>(h_steps));
  if (h_pub) (void)hipHostFree(const_cast<int*>(h_pub));
  if (h_stage) (void)hipHostFree(h_stage);
  if (ev_seg) (void)hipEventDestroy(ev_seg);
  for (auto& e : ev_pool)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

// Timing events record with a device-scope release: a system-scope one (the
// default) adds its cache write-back to the interval it closes.
// pending timed sections before their times are read back; each holds two events
constexpr size_t MAX_PENDING = 512;

static void grow_events(Ctx& c, size_t n) {
  while (c.ev_pool.size() < n) {
    hipEvent_t e = nullptr;
    CAC_HIP(hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
    c.ev_pool.push_back(e);
  }
}

static hipEvent_t take_event(Ctx& c) {
  if (c.ev_used == c.ev_pool.size()) grow_events(c, c.ev_used + 1);
  return c.ev_pool[c.ev_used++];
}

void resolve_timers(Ctx& c) {
  if (c.ev_pending.empty()) {
    c.ev_used = 0;
    return;
  }
  CAC_HIP(hipStreamSynchronize(c.stream));
  for (const auto& p : c.ev_pending) {
    float ms = 0.0f;
    CAC_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    Timer& t = c.timers[p.name];
    t.sum_us += 1000.0 * ms;
    t.count++;
  }
  c.ev_pending.clear();
  c.ev_used = 0;
}

static bool timed_section(const Ctx& c, const char* nm) {
  if (c.timing == 1) return true;
  return c.timing == 2 && (std::strcmp(nm, "eam_force") == 0 || std::strcmp(nm, "lj_force") == 0);
}

ScopedTimer::ScopedTimer(Ctx& cc, const char* nm, bool kernel_bound) : c(cc), name(nm), bound(kernel_bound) {
  if (!timed_section(c, nm)) return;
  if (c.ev_pending.size() >= MAX_PENDING) resolve_timers(c);
  start = take_event(c);
  stop = take_event(c);
  if (!bound) CAC_HIP(hipEventRecord(start, c.stream));
}

ScopedTimer::~ScopedTimer() {
  if (!start) return;
  if (!bound && hipEventRecord(stop, c.stream) != hipSuccess) return;
  c.ev_pending.push_back({name, start, stop});
}

// One velocity-Verlet step (verlet.cpp:218-360), in two halves so that a
// caller can queue the first half of several models before it waits on any.
// step_issue advances the step counter and runs Neighbor::decide: when a
// rebuild is due without a check it does the whole step and returns false.
// Otherwise it queues kick + drift + both displacement checks and the site
// refresh as if no rebuild were due (the refresh returns at once if the check
// asks for a rebuild, and refreshes every site if it asks for a re-prune), and
// returns true: step_resolve then waits for the check, rebuilds or re-prunes
// if asked, and queues the force passes and the closing half kick.
bool step_issue(Ctx& c, double dt, int ev) {
  c.ntimestep++;
  const int N = (int)c.ntimestep;
  c.ago++;
  const bool due = c.ago >= c.delay && c.ago % c.every == 0;
  const bool begun = c.begun_step == N;  // the previous step's fused pass did the opening half
  c.begun_step = -1;
  if (due && !c.check) {
    if (!begun) nve_initial(c, dt);
    build_neighbor(c);
    pair_passes(c, ev, ev, Gate());
    step_end(c, dt, Gate());
    return false;
  }
  const Gate g{c.d_steps.get(), N};
  if (!begun) step_begin(c, dt, N, due);
  site_refresh(c, false, g);
  return true;
}

// Poll a host-mapped word the device writes until it holds `want`.  After
// 1 ms the stream is also queried (then every 0.1 ms), so a failed launch is
// an error, not an endless wait; before that the poll touches no runtime
// call, which would take the runtime's locks.
static void poll_word(Ctx& c, volatile int* w, int want, const char* what) {
  using clk = std::chrono::steady_clock;
  if (*w == want) return;
  auto next_query = clk::now() + std::chrono::milliseconds(1);
  for (unsigned spin = 1; *w != want; spin++) {
    if ((spin & 1023u) != 0) continue;
    const auto t = clk::now();
    if (t < next_query) continue;
    next_query = t + std::chrono::microseconds(100);
    const hipError_t q = hipStreamQuery(c.stream);
    if (q != hipSuccess && q != hipErrorNotReady) throw HipError{q, what};
    if (q == hipSuccess && *w != want) throw HipError{hipErrorUnknown, std::string(what) + ": the device never wrote"};
  }
}

// the host-mapped word the check's last workgroup writes
void wait_check(Ctx& c) { poll_word(c, c.h_steps + 2, (int)c.ntimestep, "displacement check"); }

// One thread copies the values, fences at system scope, then writes the
// sequence number the host polls for.
__global__ void k_publish_ints(const int* __restrict__ src, int n, const int* __restrict__ a,
                               const int* __restrict__ b, int* pub, int seq) {
  if (threadIdx.x != 0) return;
  if (a)
    pub[1] = *a + *b;
  else
    for (int i = 0; i < n; i++) pub[1 + i] = src[i];
  __threadfence_system();
  *reinterpret_cast<volatile int*>(pub) = seq;
}

void read_device_ints(Ctx& c, const int* src, int n, int* dst) {
  if (n < 1 || n > 8) throw HipError{hipErrorInvalidValue, "read_device_ints: 1 to 8 values"};
  const int seq = ++c.pub_seq;
  k_publish_ints<<<1, 64, 0, c.stream>>>(src, n, nullptr, nullptr, c.h_pub_dev, seq);
  CAC_HIP(hipGetLastError());
  poll_word(c, c.h_pub, seq, "device values");
  for (int i = 0; i < n; i++) dst[i] = c.h_pub[1 + i];
}

int read_device_sum(Ctx& c, const int* a, const int* b) {
  const int seq = ++c.pub_seq;
  k_publish_ints<<<1, 64, 0, c.stream>>>(nullptr, 0, a, b, c.h_pub_dev, seq);
  CAC_HIP(hipGetLastError());
  poll_word(c, c.h_pub, seq, "device count");
  return c.h_pub[1];
}

// step_resolve in three parts, so that a caller can run the pair passes of
// several models together between the first and the last
void step_prepare(Ctx& c) {
  const int N = (int)c.ntimestep;
  wait_check(c);
  if (c.h_steps[0] == N) {
    if (c.ago == std::max(c.every, c.delay)) c.ndanger++;
    build_neighbor(c);
  } else if (c.h_steps[1] == N) {
    build_inner(c);
  }
  ensure_inner(c);
}

void step_close(Ctx& c, double dt, int ev, bool fuse_next) {
  const int N = (int)c.ntimestep;
  // Neighbor::decide of the next step, as step_issue will form it
  const bool due_next = c.ago + 1 >= c.delay && (c.ago + 1) % c.every == 0;
  if (fuse_next && !ev && c.check)
    step_end_begin(c, dt, N + 1, due_next);
  else
    step_end(c, dt, Gate());
}

void step_resolve(Ctx& c, double dt, int ev, bool fuse_next) {
  step_prepare(c);
  pair_passes(c, ev, ev, Gate());
  step_close(c, dt, ev, fuse_next);
}

// the unit constants of `units` style lj or metal (update.cpp:112-184) and
// its default neighbor skin
void set_units(Ctx& c, int units) {
  c.units = units;
  if (units == CAC_HIP_UNITS_LJ) {
    c.boltz = 1.0;
    c.mvv2e = 1.0;
    c.ftm2v = 1.0;
    c.nktv2p = 1.0;
    c.skin = 0.3;
  } else {
    c.boltz = 8.617343e-5;
    c.mvv2e = 1.0364269e-4;
    c.ftm2v = 1.0 / 1.0364269e-4;
    c.nktv2p = 1.6021765e6;
    c.skin = 2.0;
  }
}

// A context about to take a new model: box, types, element type, staged
// arrays, run counters and status words go back to their initial state; the
// pair style, the neighbor settings and every device buffer are kept, so a
// model of at most the same size loads without a device allocation.
void reset_model(Ctx& c) {
  c.box_set = false;
  c.ntypes = 0;
  c.etype_set = false;
  c.st_ax.clear();
  c.st_av.clear();
  c.st_nx.clear();
  c.st_nv.clear();
  c.st_at.clear();
  c.st_ct.clear();
  c.A = c.E = c.P = 0;
  c.model_set = false;
  c.built = false;
  c.ago = c.ncalls = c.ndanger = 0;
  c.ntimestep = 0;
  c.begun_step = -1;
  c.inner_builds = 0;
  c.inner_mode = -1;
  c.inner_due = 0;
  c.x_dirty = false;
  c.entries = -1;
  c.eng_vdwl = 0.0;
  for (double& v : c.virial) v = 0.0;
  c.peratom_flags = 0;
  c.timers.clear();
  CAC_HIP(hipMemsetAsync(c.d_flags.get(), 0, sizeof(ErrorFlags), c.stream));
  CAC_HIP(hipMemsetAsync(c.d_steps.get(), 0xFF, sizeof(int) * 2, c.stream));
  CAC_HIP(hipStreamSynchronize(c.stream));
  c.h_steps[0] = c.h_steps[1] = -1;
  c.h_steps[2] = -1;  // no check of the new model has published yet
}

}  // namespace cac

namespace {

int fail(cac_hip_ctx* h, int code, const std::string& msg) {
  if (h) h->c.err = msg;
  return code;
}

template <typename F>
int guarded(cac_hip_ctx* h, F&& fn) {
  if (!h) return CAC_HIP_EARG;
  try {
    return fn(h->c);
  } catch (const HipError& e) {
    if (e.code == hipErrorInvalidValue) {
      h->c.err = e.where;
      if (e.where.find("overflow") != std::string::npos) return CAC_HIP_EOVERFLOW;
      if (e.where.find("Non-numeric") != std::string::npos) return CAC_HIP_ENAN;
      if (e.where.find("deformed") != std::string::npos) return CAC_HIP_EDEFORM;
      if (e.where.find("larger than the periodic box") != std::string::npos) return CAC_HIP_EDEFORM;
      return CAC_HIP_EARG;
    }
    if (e.code == hipErrorOutOfMemory) {
      h->c.err = "device allocation failed: " + e.where;
      return CAC_HIP_EOOM;
    }
    h->c.err = std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where;
    return CAC_HIP_EHIP;
  } catch (const std::exception& e) {
    h->c.err = e.what();
    return CAC_HIP_EARG;
  }
}

void hex_shape(double* s, double px, double py, double pz) {  // element.cpp:2989-2997
  s[0] = 0.125 * (1 - px) * (1 - py) * (1 - pz);
  s[1] = 0.125 * (1 + px) * (1 - py) * (1 - pz);
  s[2] = 0.125 * (1 + px) * (1 + py) * (1 - pz);
  s[3] = 0.125 * (1 - px) * (1 + py) * (1 - pz);
  s[4] = 0.125 * (1 - px) * (1 - py) * (1 + pz);
  s[5] = 0.125 * (1 + px) * (1 - py) * (1 + pz);
  s[6] = 0.125 * (1 + px) * (1 + py) * (1 + pz);
  s[7] = 0.125 * (1 - px) * (1 + py) * (1 + pz);
}

const int HEX_SGN[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

template <typename T>
void upload(DBuf<T>& b, const std::vector<T>& h, hipStream_t s) {
  b.reserve(h.size());
  CAC_HIP(hipMemcpyAsync(b.get(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
  CAC_HIP(hipStreamSynchronize(s));
}

void neighbor_init(Ctx& c) {  // neighbor.cpp:241-283 with one cutoff
  c.cutforce = c.pairstyle == 1 ? std::sqrt(c.ljcutsq) : c.cutmax;
  c.cutneigh = c.cutforce + (c.cutforce > 0.0 ? c.skin : 0.0);
  c.cutneighsq = c.cutneigh * c.cutneigh;
  c.triggersq = 0.25 * c.skin * c.skin;
  c.inner_skin = c.inner_skin_set < 0.0 ? 0.3 * c.skin : std::min(c.inner_skin_set, c.skin);
}

int require_model(Ctx& c) {
  if (!c.box_set) {
    c.err = "box not set";
    return CAC_HIP_ESTATE;
  }
  if (!c.model_set) {
    c.err = "no atoms or elements uploaded";
    return CAC_HIP_ESTATE;
  }
  if (c.pairstyle == 0) {
    c.err = "no pair style";
    return CAC_HIP_ESTATE;
  }
  return CAC_HIP_OK;
}

// pinned host staging for the per-step handover, grown on demand
double* host_staging(Ctx& c, size_t n) {
  if (n > c.h_stage_n) {
    if (c.h_stage) CAC_HIP(hipHostFree(c.h_stage));
    c.h_stage = nullptr;
    CAC_HIP(hipHostMalloc((void**)&c.h_stage, sizeof(double) * n, hipHostMallocDefault));
    c.h_stage_n = n;
  }
  return c.h_stage;
}

void compute_forces(Ctx& c, int eflag, int vflag) {
  // the forward comm of a step without rebuild: refresh the sites, and re-prune
  // the inner list if a point has moved half its buffer
  if (c.x_dirty) {
    int inner = 0;
    check_distance(c, false, &inner);
    site_refresh(c, inner != 0);
    if (inner) build_inner(c);
  }
  // force_clear is implicit: every row writes its own force and every element
  // its own nodal forces
  pair_compute(c, eflag, vflag);
}

void read_error_flags(Ctx& c) {
  CAC_HIP(hipMemcpyAsync(c.h_flags, c.d_flags.get(), sizeof(ErrorFlags), hipMemcpyDeviceToHost, c.stream));
  CAC_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace

extern "C" {

int cac_hip_create(cac_hip_ctx** out, int device, int units) {
  if (!out) return CAC_HIP_EARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CAC_HIP_ENODEV;
  if (device < 0 || device >= ndev) return CAC_HIP_EARG;
  cac_hip_ctx* h = new cac_hip_ctx();
  int rc = guarded(h, [&](Ctx& c) -> int {
    c.device = device;
    CAC_HIP(hipSetDevice(device));
    CAC_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.scratch = scratch_on(device);
    CAC_HIP(hipHostMalloc((void**)&c.h_red, sizeof(double) * 64, hipHostMallocDefault));
    CAC_HIP(hipHostMalloc((void**)&c.h_flags, sizeof(ErrorFlags), hipHostMallocDefault));
    c.d_flags.reserve(1);
    CAC_HIP(hipMemsetAsync(c.d_flags.get(), 0, sizeof(ErrorFlags), c.stream));
    // step words of the speculative step: device copy for the gated kernels,
    // host-mapped copy for the host
    c.d_steps.reserve(2);

    CAC_HIP(hipMemsetAsync(c.d_steps.get(), 0xFF, sizeof(int) * 2, c.stream));
    int* hs = nullptr;
    CAC_HIP(hipHostMalloc((void**)&hs, sizeof(int) * 4, hipHostMallocMapped | hipHostMallocCoherent));
    hs[0] = hs[1] = -1;
    c.h_steps = hs;
    CAC_HIP(hipHostGetDevicePointer((void**)&c.h_steps_dev, hs, 0));
    int* hp = nullptr;
    CAC_HIP(hipHostMalloc((void**)&hp, sizeof(int) * 16, hipHostMallocMapped | hipHostMallocCoherent));
    for (int i = 0; i < 16; i++) hp[i] = 0;
    c.h_pub = hp;
    CAC_HIP(hipHostGetDevicePointer((void**)&c.h_pub_dev, hp, 0));
    c.d_check_count.reserve(9 * 64);
    CAC_HIP(hipMemsetAsync(c.d_check_count.get(), 0, sizeof(unsigned) * 9 * 64, c.stream));
    hs[2] = -1;
    CAC_HIP(hipEventCreateWithFlags(&c.ev_seg, hipEventDisableTiming));
    c.units = units;
    if (const char* e = std::getenv("CAC_HIP_TPR")) c.tpr = std::atoi(e);
    if (const char* e = std::getenv("CAC_HIP_TPR_FORCE")) c.tpr_force = std::atoi(e);
    if (const char* e = std::getenv("CAC_HIP_GRID")) c.grid_cap = std::max(8, std::atoi(e));
    if (const char* e = std::getenv("CAC_HIP_PAIR_SHAPE")) c.pair_shape = std::atoi(e);
    if (const char* e = std::getenv("CAC_HIP_DENS_NT")) c.dens_nt = std::atoi(e) == 1024 ? 1024 : 512;
    CAC_HIP(hipDeviceGetAttribute(&c.ncu, hipDeviceAttributeMultiprocessorCount, device));
    set_units(c, units);
    for (int a = 0; a < 8; a++)
      for (int b = 0; b < 8; b++) {
        int same = 0;
        for (int d = 0; d < 3; d++) same += HEX_SGN[a][d] == HEX_SGN[b][d];
        c.mmat[a][b] = same == 3 ? 1.0 / 27.0 : same == 2 ? 1.0 / 54.0 : same == 1 ? 1.0 / 108.0 : 1.0 / 216.0;
        c.minv[a][b] = same == 3 ? 64.0 : same == 2 ? -32.0 : same == 1 ? 16.0 : -8.0;
      }
    return CAC_HIP_OK;
  });
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return CAC_HIP_OK;
}

int cac_hip_destroy(cac_hip_ctx* h) {
  if (!h) return CAC_HIP_EARG;
  (void)hipSetDevice(h->c.device);
  (void)hipStreamSynchronize(h->c.stream);
  delete h;
  return CAC_HIP_OK;
}

const char* cac_hip_last_error(const cac_hip_ctx* h) { return h ? h->c.err.c_str() : "null context"; }

int cac_hip_set_box(cac_hip_ctx* h, const double* lo, const double* hi) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!lo || !hi) return fail(h, CAC_HIP_EARG, "null box");
    for (int d = 0; d < 3; d++) {
      if (!(hi[d] > lo[d])) return fail(h, CAC_HIP_EARG, "box hi must exceed lo");
      c.lo[d] = lo[d];
      c.hi[d] = hi[d];
      c.L[d] = hi[d] - lo[d];
    }
    c.box_set = true;
    return CAC_HIP_OK;
  });
}

int cac_hip_set_masses(cac_hip_ctx* h, int ntypes, const double* mass) {
  return guarded(h, [&](Ctx& c) -> int {
    if (ntypes < 1 || ntypes > MAXT || !mass) return fail(h, CAC_HIP_EARG, "ntypes out of range");
    c.ntypes = ntypes;
    for (int i = 1; i <= ntypes; i++) {
      if (!(mass[i - 1] > 0.0)) return fail(h, CAC_HIP_EARG, "mass must be positive");
      c.mass[i] = mass[i - 1];
    }
    return CAC_HIP_OK;
  });
}

// ElementVecCAC::set_element_types for Hex, apc = 1, 2x2x2 Gauss cells with the
// UNIFORM weight scale (element_vec_cac.cpp:1446-1741, 1881-1968, 2401-2558,
// 3240-3296, 3391-3468)
int cac_hip_add_etype(cac_hip_ctx* h, int itype, int ncx, int ncy, int ncz, int ngx, int ngy, int ngz) {
  return guarded(h, [&](Ctx& c) -> int {
    if (itype != 1) return fail(h, CAC_HIP_EARG, "one element type (itype 1) is supported");
    if (ncx < 2 || ncy < 2 || ncz < 2) return fail(h, CAC_HIP_EARG, "Invalid interpolate values");
    ngx = std::min(ngx, ncx);
    ngy = std::min(ngy, ncy);
    ngz = std::min(ngz, ncz);
    if (ngx != 2 || ngy != 2 || ngz != 2)
      return fail(h, CAC_HIP_EARG, "Gauss cells must be 2x2x2 (the element corners)");
    c.ncell[0] = ncx;
    c.ncell[1] = ncy;
    c.ncell[2] = ncz;
    c.nucell = ncx * ncy * ncz;
    c.nodal_weight = (double)c.nucell / (double)NPE;
    c.shape_h.assign((size_t)c.nucell * NPE, 0.0);
    c.u2g_h.assign(c.nucell, -1);
    const double dx = 2.0 / (ncx - 1), dy = 2.0 / (ncy - 1), dz = 2.0 / (ncz - 1);
    int n = 0;
    for (int i = 0; i < ncx; i++) {
      const double px = dx * i - 1.0;
      for (int j = 0; j < ncy; j++) {
        const double py = dy * j - 1.0;
        for (int k = 0; k < ncz; k++) {
          const double pz = dz * k - 1.0;
          hex_shape(&c.shape_h[(size_t)n * NPE], px, py, pz);
          n++;
        }
      }
    }
    const double w_node = (double)c.nucell / (1.0 * 8 + 0.0) * 1.0;
    int g = 0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int k = 0; k < 2; k++) {
          const int px = i == 0 ? 0 : ncx - 1, py = j == 0 ? 0 : ncy - 1, pz = k == 0 ? 0 : ncz - 1;
          const int u = px * ncy * ncz + py * ncz + pz;
          c.g2u[g] = u;
          c.u2g_h[u] = g;
          g++;
        }
    for (int q = 0; q < 8; q++) c.n2g[q] = -1;
    for (int q = 0; q < 8; q++) {
      c.g2n[q] = -1;
      const int u = c.g2u[q];
      for (int k = 0; k < NPE; k++) {
        c.wsa[q][k] = c.shape_h[(size_t)u * NPE + k] * w_node / c.nucell;
        c.wsa_lumped[q][k] = c.shape_h[(size_t)u * NPE + k] * w_node / c.nodal_weight;
        if (std::fabs(c.shape_h[(size_t)u * NPE + k] - 1.0) < 1.0e-6) {
          if (c.n2g[k] >= 0) return fail(h, CAC_HIP_EARG, "Node to gaussian cell mapping is not correct");
          c.n2g[k] = q;
          c.g2n[q] = k;
        }
      }
    }
    for (int k = 0; k < 8; k++)
      if (c.n2g[k] < 0) return fail(h, CAC_HIP_EARG, "Node to gaussian cell mapping is not complete");
    upload(c.shape, c.shape_h, c.stream);
    upload(c.u2g, c.u2g_h, c.stream);
    c.etype_set = true;
    return CAC_HIP_OK;
  });
}

}  // extern "C"

namespace {

// lay the staged atoms and element nodes out as the device point arrays
void finalize_points(Ctx& c) {
  const int A = (int)c.st_at.size(), E = (int)c.st_ct.size();
  const int P = A + NPE * E;
  std::vector<double> hx((size_t)3 * P), hv((size_t)3 * P, 0.0);
  std::vector<int> ht(P);
  for (int i = 0; i < A; i++) {
    for (int d = 0; d < 3; d++) {
      hx[(size_t)d * P + i] = c.st_ax[3 * (size_t)i + d];
      hv[(size_t)d * P + i] = c.st_av[3 * (size_t)i + d];
    }
    ht[i] = c.st_at[i];
  }
  for (int e = 0; e < E; e++)
    for (int k = 0; k < NPE; k++) {
      const size_t p = (size_t)A + (size_t)NPE * e + k;
      for (int d = 0; d < 3; d++) {
        hx[(size_t)d * P + p] = c.st_nx[((size_t)e * NPE + k) * 3 + d];
        hv[(size_t)d * P + p] = c.st_nv[((size_t)e * NPE + k) * 3 + d];
      }
      ht[p] = c.st_ct[e];
    }
  c.A = A;
  c.E = E;
  c.P = P;
  upload(c.x, hx, c.stream);
  upload(c.v, hv, c.stream);
  upload(c.ptype, ht, c.stream);
  c.f.reserve((size_t)3 * P);
  c.xhold.reserve((size_t)3 * P);
  c.xhold_in.reserve((size_t)3 * P);
  c.rho.reserve(P);
  CAC_HIP(hipMemsetAsync(c.f.get(), 0, sizeof(double) * 3 * P, c.stream));
  if (E) {
    c.gf.reserve((size_t)3 * NPE * E);
    CAC_HIP(hipMemsetAsync(c.gf.get(), 0, sizeof(double) * 3 * NPE * E, c.stream));
    c.ex.reserve((size_t)3 * E);
    c.initsize.reserve((size_t)3 * E);
    update_centers(c);
    check_deformation(c, true);  // Element::init records the initial bound boxes
  }
  CAC_HIP(hipStreamSynchronize(c.stream));
  c.model_set = P > 0;
  c.built = false;
}

}  // namespace

extern "C" {

int cac_hip_upload_atoms(cac_hip_ctx* h, int n, const double* x, const double* v, const int* type) {
  return guarded(h, [&](Ctx& c) -> int {
    if (n < 0 || (n > 0 && (!x || !type))) return fail(h, CAC_HIP_EARG, "bad atom arrays");
    if (c.ntypes == 0) return fail(h, CAC_HIP_ESTATE, "set masses before uploading atoms");
    for (int i = 0; i < n; i++)
      if (type[i] < 1 || type[i] > c.ntypes) return fail(h, CAC_HIP_EARG, "atom type out of range");
    c.st_ax.assign(x, x + 3 * (size_t)n);
    c.st_av.assign((size_t)3 * n, 0.0);
    if (v) c.st_av.assign(v, v + 3 * (size_t)n);
    c.st_at.assign(type, type + n);
    finalize_points(c);
    return CAC_HIP_OK;
  });
}

int cac_hip_upload_elements(cac_hip_ctx* h, int n, const int* etype, const int* ctype, const double* nodex,
                            const double* nodev) {
  return guarded(h, [&](Ctx& c) -> int {
    if (n < 0 || (n > 0 && (!etype || !ctype || !nodex))) return fail(h, CAC_HIP_EARG, "bad element arrays");
    if (n > 0 && !c.etype_set) return fail(h, CAC_HIP_ESTATE, "define the element type first");
    if (c.ntypes == 0) return fail(h, CAC_HIP_ESTATE, "set masses before uploading elements");
    for (int i = 0; i < n; i++) {
      if (etype[i] != 1) return fail(h, CAC_HIP_EARG, "element type out of range");
      if (ctype[i] < 1 || ctype[i] > c.ntypes) return fail(h, CAC_HIP_EARG, "element ctype out of range");
    }
    const size_t m = (size_t)n * NPE * 3;
    c.st_nx.assign(nodex, nodex + m);
    c.st_nv.assign(m, 0.0);
    if (nodev) c.st_nv.assign(nodev, nodev + m);
    c.st_ct.assign(ctype, ctype + n);
    finalize_points(c);
    return CAC_HIP_OK;
  });
}

int cac_hip_pair_lj(cac_hip_ctx* h, double eps, double sigma, double cut) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!(cut > 0.0) || !(sigma > 0.0)) return fail(h, CAC_HIP_EARG, "bad lj/cut coefficients");
    c.pairstyle = 1;
    c.lj1 = 48.0 * eps * std::pow(sigma, 12.0);
    c.lj2 = 24.0 * eps * std::pow(sigma, 6.0);
    c.lj3 = 4.0 * eps * std::pow(sigma, 12.0);
    c.lj4 = 4.0 * eps * std::pow(sigma, 6.0);
    c.ljcutsq = cut * cut;
    c.built = false;
    return CAC_HIP_OK;
  });
}

int cac_hip_pair_eam(cac_hip_ctx* h, int nrho, double drho, int nr, double dr, double cut, const double* frho,
                     const double* rhor, const double* z2r) {
  return guarded(h, [&](Ctx& c) -> int {
    if (nrho < 5 || nr < 5 || !(dr > 0.0) || !(drho > 0.0) || !frho || !rhor || !z2r)
      return fail(h, CAC_HIP_EARG, "Invalid EAM potential file");
    c.pairstyle = 2;
    c.nrho = nrho;
    c.nr = nr;
    c.drho = drho;
    c.dr = dr;
    c.rdr = 1.0 / dr;
    c.rdrho = 1.0 / drho;
    c.rhomax = (nrho - 1) * drho;
    c.cutmax = cut;
    c.cutforcesq = cut * cut;
    std::vector<double> t(std::max(nrho, nr) + 1, 0.0), fs, rs, zs;
    for (int m = 1; m <= nrho; m++) t[m] = frho[m - 1];
    interpolate(nrho, drho, t, fs);
    for (int m = 1; m <= nr; m++) t[m] = rhor[m - 1];
    interpolate(nr, dr, t, rs);
    for (int m = 1; m <= nr; m++) t[m] = z2r[m - 1];
    interpolate(nr, dr, t, zs);
    upload(c.frho_spl, fs, c.stream);
    upload(c.rhor_spl, rs, c.stream);
    upload(c.z2r_spl, zs, c.stream);
    // FP32 tables of the mixed path, staged in LDS by the pair kernels: the
    // density pass reads rhor's value coefficients c3..c6 per knot, the force
    // pass the planes A, B (and C) of mixed_force_table.  Like the FP64
    // tables, one knot more than the reference's: knot nr + 1 is all zeros,
    // and the pair passes evaluate it for a list entry beyond the cutoff, so
    // such an entry contributes an exact zero without a mask multiply.
    std::vector<float> t32r((size_t)(nr + 2) * 4, 0.0f), t32f;
    for (int m = 0; m <= nr; m++)
      for (int q = 0; q < 4; q++) t32r[(size_t)m * 4 + q] = (float)rs[(size_t)m * 7 + 3 + q];
    mixed_force_table(nr, dr, rs, zs, t32f);
    std::vector<double> t64((size_t)(nr + 2) * 8, 0.0);
    for (int m = 0; m <= nr; m++)
      for (int q = 0; q < 4; q++) {
        t64[(size_t)m * 8 + q] = rs[(size_t)m * 7 + 3 + q];
        t64[(size_t)m * 8 + 4 + q] = zs[(size_t)m * 7 + 3 + q];
      }
    if ((size_t)(nr + 2) * 64 > 160 * 1024) return fail(h, CAC_HIP_EARG, "nr too large for the LDS spline table");
    upload(c.tab32r, t32r, c.stream);
    upload(c.tab32f, t32f, c.stream);
    upload(c.tab64, t64, c.stream);
    c.built = false;
    return CAC_HIP_OK;
  });
}

// PairEAMAlloy::read_file for a one-element setfl (pair_eam_alloy.cpp:100-196)
int cac_hip_pair_eam_setfl(cac_hip_ctx* h, const char* path) {
  if (!h || !path) return CAC_HIP_EARG;
  std::ifstream in(path);
  if (!in) return fail(h, CAC_HIP_EARG, std::string("Cannot open EAM potential file ") + path);
  std::string line;
  for (int i = 0; i < 3; i++) std::getline(in, line);
  std::getline(in, line);
  int nelem = std::atoi(line.c_str());
  if (nelem != 1) return fail(h, CAC_HIP_EARG, "one-element setfl files are read");
  std::getline(in, line);
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, cut = 0;
  {
    std::istringstream ss(line);
    if (!(ss >> nrho >> drho >> nr >> dr >> cut)) return fail(h, CAC_HIP_EARG, "Invalid EAM potential file");
  }
  std::getline(in, line);  // Z mass a lattice
  std::vector<double> frho(nrho), rhor(nr), z2r(nr);
  for (auto* vec : {&frho, &rhor, &z2r})
    for (auto& x : *vec)
      if (!(in >> x)) return fail(h, CAC_HIP_EARG, "truncated setfl file");
  return cac_hip_pair_eam(h, nrho, drho, nr, dr, cut, frho.data(), rhor.data(), z2r.data());
}

int cac_hip_set_precision(cac_hip_ctx* h, int mode) {
  return guarded(h, [&](Ctx& c) -> int {
    if (mode != CAC_HIP_FP64 && mode != CAC_HIP_MIXED) return fail(h, CAC_HIP_EARG, "precision must be 0 or 1");
    c.precision = mode;
    return CAC_HIP_OK;
  });
}

int cac_hip_set_nodal_force_style(cac_hip_ctx* h, int style) {
  return guarded(h, [&](Ctx& c) -> int {
    if (style < 0 || style > 2) return fail(h, CAC_HIP_EARG, "nodal force style must be 0, 1 or 2");
    c.nodef_style = style;
    return CAC_HIP_OK;
  });
}

int cac_hip_neighbor_settings(cac_hip_ctx* h, double skin, int every, int delay, int check, int one) {
  return guarded(h, [&](Ctx& c) -> int {
    if (skin < 0.0 || every < 1 || delay < 0) return fail(h, CAC_HIP_EARG, "bad neighbor settings");
    if (delay > 0 && delay % every) return fail(h, CAC_HIP_EARG, "Neighbor delay must be 0 or multiple of every setting");
    c.skin = skin;
    c.every = every;
    c.delay = delay;
    c.check = check ? 1 : 0;
    if (one > 0) c.one = 2 * one;  // full rows hold both halves of a half row
    c.built = false;
    return CAC_HIP_OK;
  });
}

// Domain::pbc (domain.cpp:526-738) on the device model, as write_data runs it
// before writing (write_data.cpp:120-131); the lists are stale afterwards
int cac_hip_pbc(cac_hip_ctx* h) {
  return guarded(h, [&](Ctx& c) -> int {
    int rc = require_model(c);
    if (rc) return rc;
    CAC_HIP(hipSetDevice(c.device));
    wrap_model(c);
    c.built = false;
    return CAC_HIP_OK;
  });
}

int cac_hip_neigh_build(cac_hip_ctx* h) {
  return guarded(h, [&](Ctx& c) -> int {
    int rc = require_model(c);
    if (rc) return rc;
    CAC_HIP(hipSetDevice(c.device));
    neighbor_init(c);
    build_neighbor(c);
    return CAC_HIP_OK;
  });
}

int cac_hip_neigh_decide(cac_hip_ctx* h, int* flag) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!c.built) return fail(h, CAC_HIP_ESTATE, "neighbor list not built");
    int f = 0, inner = 0;
    c.ago++;
    const bool due = c.ago >= c.delay && c.ago % c.every == 0;
    if (due && !c.check)
      f = 1;
    else {
      f = check_distance(c, due && c.check, &inner);
      if (f && c.ago == std::max(c.every, c.delay)) c.ndanger++;
    }
    c.inner_due = inner;
    if (flag) *flag = f;
    return CAC_HIP_OK;
  });
}

static long count_entries(Ctx& c, const DBuf<int>& nnum) {
  if (!c.built || !c.P) return 0;
  std::vector<int> h(c.P);
  CAC_HIP(hipStreamSynchronize(c.stream));
  CAC_HIP(hipMemcpy(h.data(), nnum.get(), sizeof(int) * c.P, hipMemcpyDeviceToHost));
  long s = 0;
  for (int v : h) s += v;
  return s;
}

int cac_hip_neigh_info(cac_hip_ctx* h, int* rows, long* entries, int* sites, int* capacity) {
  return guarded(h, [&](Ctx& c) -> int {
    if (c.entries < 0) c.entries = count_entries(c, c.nnum);
    if (rows) *rows = c.P;
    if (entries) *entries = c.entries;
    if (sites) *sites = c.S;
    if (capacity) *capacity = c.cap;
    return CAC_HIP_OK;
  });
}

int cac_hip_neigh_export(cac_hip_ctx* h, long* n, int* row, int* kind, int* owner, int* ucell, int* img) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!c.built) return fail(h, CAC_HIP_ESTATE, "neighbor list not built");
    if (!n) return fail(h, CAC_HIP_EARG, "null count");
    if (!row) {
      if (c.entries < 0) c.entries = count_entries(c, c.nnum);
      *n = c.entries;
      return CAC_HIP_OK;
    }
    std::vector<int> r, k, o, u, im;
    export_lists(c, r, k, o, u, im);
    const long m = (long)r.size();
    if (*n < m) return fail(h, CAC_HIP_EARG, "export buffer too small");
    std::memcpy(row, r.data(), sizeof(int) * m);
    std::memcpy(kind, k.data(), sizeof(int) * m);
    std::memcpy(owner, o.data(), sizeof(int) * m);
    std::memcpy(ucell, u.data(), sizeof(int) * m);
    std::memcpy(img, im.data(), sizeof(int) * 3 * m);
    *n = m;
    return CAC_HIP_OK;
  });
}

int cac_hip_compute(cac_hip_ctx* h, int eflag, int vflag) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!c.built) return fail(h, CAC_HIP_ESTATE, "neighbor list not built");
    if (eflag < 0 || eflag > 3 || vflag < 0 || vflag > 15)
      return fail(h, CAC_HIP_EARG, "eflag must be 0..3 and vflag 0..15 (force.h:11-19)");
    CAC_HIP(hipSetDevice(c.device));
    compute_forces(c, eflag, vflag);
    CAC_HIP(hipStreamSynchronize(c.stream));
    return CAC_HIP_OK;
  });
}

// Pair::eatom / enode / vatom / vnode of the last force computation
// (pair.h:13-17): per point, ordered as the caller's arrays, from the [7][P]
// per-point rows the force pass wrote (point A + 8e + k is node k of element e)
int cac_hip_get_peratom(cac_hip_ctx* h, double* eatom, double* enode, double* vatom, double* vnode) {
  return guarded(h, [&](Ctx& c) -> int {
    if ((eatom || enode) && !(c.peratom_flags & 1))
      return fail(h, CAC_HIP_ESTATE, "per-atom energy was not computed (eflag & ENERGY_ATOM)");
    if ((vatom || vnode) && !(c.peratom_flags & 2))
      return fail(h, CAC_HIP_ESTATE, "per-atom virial was not computed (vflag & VIRIAL_ATOM or VIRIAL_CENTROID)");
    const int A = c.A, P = c.P;
    const size_t E8 = (size_t)NPE * c.E;
    const bool want_v = vatom || vnode;
    std::vector<double> b((size_t)(want_v ? 7 : 1) * P);
    CAC_HIP(hipStreamSynchronize(c.stream));
    if (P) CAC_HIP(hipMemcpy(b.data(), c.peratom.get(), sizeof(double) * b.size(), hipMemcpyDeviceToHost));
    if (eatom) std::memcpy(eatom, b.data(), sizeof(double) * A);
    if (enode) std::memcpy(enode, b.data() + A, sizeof(double) * E8);
    for (int k = 0; k < 6 && want_v; k++) {
      const double* src = b.data() + (size_t)(k + 1) * P;
      if (vatom)
        for (int i = 0; i < A; i++) vatom[6 * (size_t)i + k] = src[i];
      if (vnode)
        for (size_t q = 0; q < E8; q++) vnode[6 * q + k] = src[A + q];
    }
    return CAC_HIP_OK;
  });
}

int cac_hip_get_energy(cac_hip_ctx* h, double* eng, double* virial) {
  return guarded(h, [&](Ctx& c) -> int {
    if (eng) *eng = c.eng_vdwl;
    if (virial)
      for (int i = 0; i < 6; i++) virial[i] = c.virial[i];
    return CAC_HIP_OK;
  });
}

int cac_hip_nve_initial(cac_hip_ctx* h, double dt) {
  return guarded(h, [&](Ctx& c) -> int {
    nve_initial(c, dt);
    return CAC_HIP_OK;
  });
}

int cac_hip_nve_final(cac_hip_ctx* h, double dt) {
  return guarded(h, [&](Ctx& c) -> int {
    nve_final(c, dt);
    return CAC_HIP_OK;
  });
}

int cac_hip_setup(cac_hip_ctx* h, int eflag, int vflag) {
  return guarded(h, [&](Ctx& c) -> int {
    int rc = require_model(c);
    if (rc) return rc;
    if (eflag < 0 || eflag > 3 || vflag < 0 || vflag > 15)
      return fail(h, CAC_HIP_EARG, "eflag must be 0..3 and vflag 0..15 (force.h:11-19)");
    CAC_HIP(hipSetDevice(c.device));
    neighbor_init(c);
    build_neighbor(c);
    c.ncalls = 0;
    compute_forces(c, eflag, vflag);
    CAC_HIP(hipStreamSynchronize(c.stream));
    return CAC_HIP_OK;
  });
}

// ComputePressure::compute_scalar (compute_pressure.cpp:140-180) from the
// temperature and the pair virial of the last global tally
static double pressure(const Ctx& c, double t) {
  const double dof = 3.0 * ((double)c.A + (double)c.nucell * c.E);
  const double inv_volume = 1.0 / (c.L[0] * c.L[1] * c.L[2]);
  return (dof * c.boltz * t + c.virial[0] + c.virial[1] + c.virial[2]) / 3.0 * inv_volume * c.nktv2p;
}

int cac_hip_thermo(cac_hip_ctx* h, double* pe, double* ke, double* temp, double* press) {
  return guarded(h, [&](Ctx& c) -> int {
    double k = 0, t = 0;
    thermo(c, &k, &t);
    if (pe) *pe = c.eng_vdwl;
    if (ke) *ke = k;
    if (temp) *temp = t;
    if (press) *press = pressure(c, t);
    return CAC_HIP_OK;
  });
}

int cac_hip_thermo_noalloc(cac_hip_ctx* h, double* pe, double* ke, double* temp, double* press) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!c.model_set) return fail(h, CAC_HIP_ESTATE, "no atoms or elements uploaded");
    CAC_HIP(hipSetDevice(c.device));
    double k = 0, t = 0;
    thermo_noalloc(c, &k, &t);
    if (pe) *pe = c.eng_vdwl;
    if (ke) *ke = k;
    if (temp) *temp = t;
    if (press) *press = pressure(c, t);
    return CAC_HIP_OK;
  });
}

// Verlet::run (verlet.cpp:218-360) with FixNVE and thermo output
int cac_hip_run(cac_hip_ctx* h, int nsteps, double dt, int thermo_every, double* th) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!c.built) return fail(h, CAC_HIP_ESTATE, "call cac_hip_setup first");
    CAC_HIP(hipSetDevice(c.device));
    int nt = 0;
    auto record = [&]() {
      double k = 0, t = 0;
      thermo(c, &k, &t);
      const double press = pressure(c, t);
      double* r = th + 5 * nt++;
      r[0] = (double)c.ntimestep;
      r[1] = c.eng_vdwl;
      r[2] = k;
      r[3] = c.eng_vdwl + k;
      r[4] = press;
    };
    if (th) record();
    c.begun_step = -1;
    for (int s = 0; s < nsteps; s++) {
      const bool last = s == nsteps - 1;
      const int ev = th && ((thermo_every > 0 && (c.ntimestep + 1) % thermo_every == 0) || last) ? 1 : 0;
      if (step_issue(c, dt, ev)) step_resolve(c, dt, ev, !last);
      if (ev) pair_sums(c, ev, ev);
      if (ev) record();
    }
    read_error_flags(c);
    if (c.h_flags->nan) return fail(h, CAC_HIP_ENAN, "Non-numeric positions - simulation unstable");
    return CAC_HIP_OK;
  });
}

int cac_hip_download_atoms(cac_hip_ctx* h, double* x, double* v, double* f) {
  return guarded(h, [&](Ctx& c) -> int {
    CAC_HIP(hipStreamSynchronize(c.stream));
    const int A = c.A, P = c.P;
    std::vector<double> b((size_t)3 * P);
    const DBuf<double>* src[3] = {&c.x, &c.v, &c.f};
    double* dst[3] = {x, v, f};
    for (int q = 0; q < 3; q++) {
      if (!dst[q] || !A) continue;
      CAC_HIP(hipMemcpy(b.data(), src[q]->get(), sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
      for (int i = 0; i < A; i++)
        for (int d = 0; d < 3; d++) dst[q][3 * (size_t)i + d] = b[(size_t)d * P + i];
    }
    return CAC_HIP_OK;
  });
}

int cac_hip_download_elements(cac_hip_ctx* h, double* nodex, double* nodev, double* nodef, double* gaussf) {
  return guarded(h, [&](Ctx& c) -> int {
    CAC_HIP(hipStreamSynchronize(c.stream));
    const int A = c.A, P = c.P, E = c.E;
    if (!E) return CAC_HIP_OK;
    std::vector<double> b((size_t)3 * P);
    const DBuf<double>* src[3] = {&c.x, &c.v, &c.f};
    double* dst[3] = {nodex, nodev, nodef};
    for (int q = 0; q < 3; q++) {
      if (!dst[q]) continue;
      CAC_HIP(hipMemcpy(b.data(), src[q]->get(), sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
      for (int e = 0; e < E; e++)
        for (int k = 0; k < NPE; k++)
          for (int d = 0; d < 3; d++)
            dst[q][((size_t)e * NPE + k) * 3 + d] = b[(size_t)d * P + A + (size_t)NPE * e + k];
    }
    if (gaussf) {
      const size_t E8 = (size_t)NPE * E;
      std::vector<double> g(3 * E8);
      CAC_HIP(hipMemcpy(g.data(), c.gf.get(), sizeof(double) * 3 * E8, hipMemcpyDeviceToHost));
      for (int e = 0; e < E; e++)
        for (int q = 0; q < 8; q++)
          for (int d = 0; d < 3; d++)
            gaussf[((size_t)e * 8 + q) * 3 + d] = g[(size_t)d * E8 + (size_t)NPE * e + c.g2n[q]];
    }
    return CAC_HIP_OK;
  });
}

int cac_hip_sync_positions(cac_hip_ctx* h, const double* x, const double* nodex) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!c.model_set) return fail(h, CAC_HIP_ESTATE, "no atoms or elements uploaded");
    if ((c.A && !x) || (c.E && !nodex)) return fail(h, CAC_HIP_EARG, "null positions");
    const int A = c.A, P = c.P, E = c.E;
    double* st = host_staging(c, (size_t)3 * P);
    CAC_HIP(hipStreamSynchronize(c.stream));  // the staging buffer may still feed an earlier copy
    for (int d = 0; d < 3; d++) {
      double* o = st + (size_t)d * P;
      for (int i = 0; i < A; i++) o[i] = x[3 * (size_t)i + d];
      for (size_t q = 0; q < (size_t)NPE * E; q++) o[A + q] = nodex[3 * q + d];
    }
    CAC_HIP(hipMemcpyAsync(c.x.get(), st, sizeof(double) * 3 * P, hipMemcpyHostToDevice, c.stream));
    c.x_dirty = true;
    return CAC_HIP_OK;
  });
}

int cac_hip_download_forces(cac_hip_ctx* h, double* f, double* nodef, double* gaussf) {
  return guarded(h, [&](Ctx& c) -> int {
    const int A = c.A, P = c.P, E = c.E;
    const size_t E8 = (size_t)NPE * E;
    double* st = host_staging(c, (size_t)3 * P + 3 * E8);
    CAC_HIP(hipMemcpyAsync(st, c.f.get(), sizeof(double) * 3 * P, hipMemcpyDeviceToHost, c.stream));
    if (gaussf && E)
      CAC_HIP(hipMemcpyAsync(st + (size_t)3 * P, c.gf.get(), sizeof(double) * 3 * E8, hipMemcpyDeviceToHost,
                             c.stream));
    CAC_HIP(hipStreamSynchronize(c.stream));
    if (f)
      for (int i = 0; i < A; i++)
        for (int d = 0; d < 3; d++) f[3 * (size_t)i + d] = st[(size_t)d * P + i];
    if (nodef)
      for (size_t q = 0; q < E8; q++)
        for (int d = 0; d < 3; d++) nodef[3 * q + d] = st[(size_t)d * P + A + q];
    if (gaussf) {
      const double* g = st + (size_t)3 * P;
      for (int e = 0; e < E; e++)
        for (int q = 0; q < 8; q++)
          for (int d = 0; d < 3; d++)
            gaussf[((size_t)e * 8 + q) * 3 + d] = g[(size_t)d * E8 + (size_t)NPE * e + c.g2n[q]];
    }
    return CAC_HIP_OK;
  });
}

int cac_hip_download_rho(cac_hip_ctx* h, double* atomrho, double* noderho) {
  return guarded(h, [&](Ctx& c) -> int {
    CAC_HIP(hipStreamSynchronize(c.stream));
    std::vector<double> r(c.P);
    CAC_HIP(hipMemcpy(r.data(), c.rho.get(), sizeof(double) * c.P, hipMemcpyDeviceToHost));
    if (atomrho) std::memcpy(atomrho, r.data(), sizeof(double) * c.A);
    if (noderho) std::memcpy(noderho, r.data() + c.A, sizeof(double) * NPE * c.E);
    return CAC_HIP_OK;
  });
}

int cac_hip_counts(cac_hip_ctx* h, int* natoms, int* nelements, int* npoints) {
  return guarded(h, [&](Ctx& c) -> int {
    if (natoms) *natoms = c.A;
    if (nelements) *nelements = c.E;
    if (npoints) *npoints = c.P;
    return CAC_HIP_OK;
  });
}

int cac_hip_sync(cac_hip_ctx* h) {
  return guarded(h, [&](Ctx& c) -> int {
    CAC_HIP(hipStreamSynchronize(c.stream));
    return CAC_HIP_OK;
  });
}

int cac_hip_enable_timing(cac_hip_ctx* h, int on) {
  return guarded(h, [&](Ctx& c) -> int {
    if (on < 0 || on > 2) return fail(h, CAC_HIP_EARG, "timing mode must be 0, 1 or 2");
    resolve_timers(c);
    c.timing = on;
    c.timers.clear();
    // every event the timed sections will take before a read-back, created
    // here: a creation costs host time that a host-bound step would show
    if (on) grow_events(c, 2 * MAX_PENDING);
    return CAC_HIP_OK;
  });
}

int cac_hip_kernel_time(cac_hip_ctx* h, const char* name, double* avg_us, long* launches) {
  return guarded(h, [&](Ctx& c) -> int {
    resolve_timers(c);
    auto it = c.timers.find(name ? name : "");
    if (it == c.timers.end()) {
      if (avg_us) *avg_us = 0.0;
      if (launches) *launches = 0;
      return CAC_HIP_OK;
    }
    if (avg_us) *avg_us = it->second.count ? it->second.sum_us / it->second.count : 0.0;
    if (launches) *launches = it->second.count;
    return CAC_HIP_OK;
  });
}

int cac_hip_neigh_stats(cac_hip_ctx* h, int* ncalls, int* ndanger) {
  return guarded(h, [&](Ctx& c) -> int {
    if (ncalls) *ncalls = c.ncalls;
    if (ndanger) *ndanger = c.ndanger;
    return CAC_HIP_OK;
  });
}

int cac_hip_inner_stats(cac_hip_ctx* h, long* builds, double* mean_entries, double* delta, int* live_sites) {
  return guarded(h, [&](Ctx& c) -> int {
    if (live_sites) {
      int n = 0;
      if (c.built && c.inner_mode >= 0) {
        CAC_HIP(hipStreamSynchronize(c.stream));
        CAC_HIP(hipMemcpy(&n, c.d_nlive.get(), sizeof(int), hipMemcpyDeviceToHost));
      }
      *live_sites = n;
    }
    if (builds) *builds = c.inner_builds;
    if (delta) *delta = c.inner_skin;
    if (mean_entries) *mean_entries = c.P && c.inner_mode >= 0 ? (double)count_entries(c, c.innum) / c.P : 0.0;
    return CAC_HIP_OK;
  });
}

int cac_hip_set_inner_skin(cac_hip_ctx* h, double delta) {
  return guarded(h, [&](Ctx& c) -> int {
    if (!(delta == delta)) return fail(h, CAC_HIP_EARG, "inner skin must be a number");
    c.inner_skin_set = delta;
    neighbor_init(c);
    c.inner_mode = -1;  // re-prune at the next force call
    return CAC_HIP_OK;
  });
}

}  // extern "C"
