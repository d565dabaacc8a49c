// serve.hip -- the resident serving loop of include/cac_hip.h (cac_hip_server_*).
//
// Independent replicas (whole models, each a Verlet::run of its own,
// verlet.cpp:218-360) are submitted from any host thread and served by
// service threads over a fixed set of slots: a loader that admits and loads
// replicas, stepping threads that step the loaded ones, and a rebuilder that
// finishes the steps that need an outer-list rebuild.  A slot is a device context
// with its own stream; it keeps its buffers from one replica to the next, so
// serving never frees device memory (a free synchronizes the whole device).
//
// Admission: replicas wait in arrival order; the head of the queue is admitted
// when a slot is free and the HBM it needs, estimated from its size, fits the
// budget beside what every slot already holds (a slot's holding only grows).
//
// Stepping: every round queues the first half of a step (kick, drift, both
// displacement checks, site refresh) for every active replica, then walks the
// replicas again, waits for each one's check, rebuilds or re-prunes if asked
// and queues the force passes and the closing kick.  By the time the stepping
// thread comes back to a replica its check has long finished, so the thread
// rarely waits, and the kernels of different replicas, on different streams,
// overlap on the device.  A replica's results do not depend on the others:
// its kernels, launch shapes and summation order are those of a lone run.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/cac_hip.h"
#include "cac_ctx.hpp"

using namespace cac;

namespace {

using Clock = std::chrono::steady_clock;

// an owned copy of a submitted request
struct Request {
  cac_hip_request r{};
  std::vector<double> mass, x, v, nodex, nodev;
  std::vector<int> type, ctype, etype;
  long id = 0;
  double t_submit = 0.0;
  double need = 0.0;  // estimated device bytes
};

struct Outcome {
  cac_hip_result r{};
  std::vector<double> x, nodex;
  bool done = false;
};

// Device-side step completion: an event recorded on the slot's stream after
// each step of its replica (and after the load), read back when it has
// completed; the time between two consecutive completions is the replica's
// step interval as the device ran it, queueing behind the other replicas'
// kernels included.  At most MARKS - 1 marks are outstanding, so the event of
// the last completion read is never recorded over before it is used.
constexpr int MARKS = 16;

struct Slot {
  cac_hip_ctx* h = nullptr;
  double held = 0.0;  // device bytes this slot's buffers are estimated to hold
  bool busy = false;
  hipEvent_t marks[MARKS] = {};
  long head = 0, tail = 0;  // marks recorded / read back
  int last = -1;            // the last completed mark of the current replica, or -1
  double dev_sum = 0.0;     // the current replica's device step intervals so far (ms)
  long dev_n = 0;
};

struct Active {
  std::unique_ptr<Request> req;
  int slot = -1;
  int stepper = -1;  // the stepping thread that owns it
  int done = 0;
  double t_admit = 0.0, t_first = 0.0, t_last = 0.0;
  bool issued = false;
  bool away = false;  // handed to the rebuilder for this step
};

// Device bytes a model needs, from its sizes: per point the two lists
// (about 230 entries of capacity each at rc + skin = 6.95 A, with headroom) and
// the point arrays; per represented site the candidate sort arrays; per atom
// and per element the compacted site records (an element keeps about 270 of
// its 9^3 sites referenced).
double need_bytes(const cac_hip_request& q) {
  const double P = (double)q.natoms + 8.0 * q.nelements;
  const double nucell = (double)q.ncell[0] * q.ncell[1] * q.ncell[2];
  const double rep = (double)q.natoms + nucell * q.nelements;
  const double outer = (double)q.natoms + 270.0 * q.nelements;
  return 2600.0 * P + 40.0 * rep + 150.0 * outer + (64 << 20);
}

}  // namespace

struct cac_hip_server {
  int device = 0;
  double budget = 0.0;
  std::string err;
  std::vector<Slot> slots;
  bool pair_set = false, submitted = false;

  mutable std::mutex mu;  // mutable: the const accessors lock it too
  std::condition_variable cv_work, cv_ready, cv_done;
  std::deque<std::unique_ptr<Request>> queue;
  std::vector<std::vector<Active>> ready;  // per stepping thread: loaded, not yet stepped
  std::vector<int> load_of;                // per stepping thread: replicas assigned
  std::map<long, Outcome> out;
  long next_id = 0, finished = 0;
  int in_flight = 0;  // admitted and not finished (loading, ready or stepping)
  bool stop = false, loader_done = false;
  std::vector<std::thread> steppers;
  std::thread loader;
  // replicas whose step needs an outer-list rebuild, taken off their stepping
  // thread (a rebuild waits on the host for counts) and stepped by the rebuilder
  std::deque<Active> rebuild_q;
  std::condition_variable cv_rebuild;
  bool rebuild_stop = false;
  std::vector<std::thread> rebuilders;
  Clock::time_point t0;

  // statistics, written by the service thread under mu
  std::vector<float> interval_ms;
  std::vector<float> device_ms;  // step completion intervals on the device
  long seg_launches = 0, seg_steps = 0;
  double peak_hbm = 0.0;
  int peak_active = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

int set_err(cac_hip_server* s, int code, const std::string& m) {
  if (s) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = m;
  }
  return code;
}

// Read back the completed marks of a slot in order; wait for all of them
// when `all`.  Each completion after the first of a replica adds the time
// since the previous one to the device intervals.
void harvest_marks(cac_hip_server* s, Slot& sl, bool all) {
  std::vector<float> got;
  while (sl.tail < sl.head) {
    hipEvent_t e = sl.marks[sl.tail % MARKS];
    if (all) {
      if (hipEventSynchronize(e) != hipSuccess) break;
    } else if (hipEventQuery(e) != hipSuccess) {
      break;
    }
    if (sl.last >= 0) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, sl.marks[sl.last], e) == hipSuccess) {
        got.push_back(ms);
        sl.dev_sum += ms;
        sl.dev_n++;
      }
    }
    sl.last = (int)(sl.tail % MARKS);
    sl.tail++;
  }
  if (!got.empty()) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->device_ms.insert(s->device_ms.end(), got.begin(), got.end());
  }
}

// a step (or the load) of the slot's replica has been queued: mark its completion
void mark_step(cac_hip_server* s, Slot& sl) {
  if (sl.head - sl.tail >= MARKS - 1) harvest_marks(s, sl, true);
  if (hipEventRecord(sl.marks[sl.head % MARKS], sl.h->c.stream) == hipSuccess) sl.head++;
  harvest_marks(s, sl, false);
}

// load a request into a slot's context and set it up (neighbor build + first
// force); returns 0 or the CAC_HIP_E* code
int load(Slot& sl, const Request& q, std::string& msg) {
  cac_hip_ctx* h = sl.h;
  Ctx& c = h->c;
  try {
    reset_model(c);
  } catch (const HipError& e) {
    msg = std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where;
    return CAC_HIP_EHIP;
  }
  const cac_hip_request& r = q.r;
  int rc = 0;
  auto step = [&](int code) {
    if (!rc && code) {
      rc = code;
      msg = cac_hip_last_error(h);
    }
  };
  if (r.units != CAC_HIP_UNITS_LJ && r.units != CAC_HIP_UNITS_METAL) {
    msg = "units must be CAC_HIP_UNITS_LJ or CAC_HIP_UNITS_METAL";
    return CAC_HIP_EARG;
  }
  set_units(c, r.units);  // a slot takes each replica's units
  step(cac_hip_set_box(h, r.lo, r.hi));
  step(cac_hip_set_masses(h, r.ntypes, r.mass));
  if (r.nelements > 0) step(cac_hip_add_etype(h, 1, r.ncell[0], r.ncell[1], r.ncell[2], 2, 2, 2));
  step(cac_hip_upload_atoms(h, r.natoms, r.x, r.v, r.type));
  step(cac_hip_upload_elements(h, r.nelements, q.etype.data(), r.ctype, r.nodex, r.nodev));
  step(cac_hip_set_precision(h, r.precision));
  step(cac_hip_neighbor_settings(h, r.skin, 1, 10, 1, 2000));
  step(cac_hip_setup(h, 1, 1));  // Verlet::setup: lists, then forces with energy and virial
  return rc;
}

// the C ABI's code for an error a launcher raised (as the guarded entry points map them)
int code_of(const HipError& e) {
  if (e.code == hipErrorInvalidValue) {
    if (e.where.find("overflow") != std::string::npos) return CAC_HIP_EOVERFLOW;
    if (e.where.find("Non-numeric") != std::string::npos) return CAC_HIP_ENAN;
    if (e.where.find("deformed") != std::string::npos) return CAC_HIP_EDEFORM;
    if (e.where.find("larger than the periodic box") != std::string::npos) return CAC_HIP_EDEFORM;
    return CAC_HIP_EARG;
  }
  return e.code == hipErrorOutOfMemory ? CAC_HIP_EOOM : CAC_HIP_EHIP;
}

void finish(cac_hip_server* s, Active& a, int status, const std::string& msg) {
  Slot& sl = s->slots[a.slot];
  Ctx& c = sl.h->c;
  harvest_marks(s, sl, true);
  sl.last = -1;
  const double dev_mean = sl.dev_n ? sl.dev_sum / (double)sl.dev_n : 0.0;
  sl.dev_sum = 0.0;
  sl.dev_n = 0;
  Outcome o;
  o.r.status = status;
  o.r.nsteps = a.req->r.nsteps;
  o.r.atom_equivalents = (long)a.req->r.natoms + 8L * a.req->r.nelements;
  o.r.t_submit = a.req->t_submit;
  o.r.t_admit = a.t_admit;
  o.r.t_done = s->now();
  o.r.step_ms_mean = a.done > 0 ? 1000.0 * (o.r.t_done - a.t_first) / a.done : 0.0;
  o.r.device_step_ms_mean = dev_mean;
  o.r.builds = c.ncalls;
  if (status == 0) {
    try {
      pair_sums(c, 1, 1);
      double ke = 0.0, t = 0.0;
      thermo(c, &ke, &t);
      o.r.pe = c.eng_vdwl;
      o.r.ke = ke;
      o.r.etotal = c.eng_vdwl + ke;
      CAC_HIP(hipMemcpyAsync(c.h_flags, c.d_flags.get(), sizeof(ErrorFlags), hipMemcpyDeviceToHost, c.stream));
      CAC_HIP(hipStreamSynchronize(c.stream));
      if (c.h_flags->nan) o.r.status = CAC_HIP_ENAN;
      if (a.req->r.keep_state) {
        o.x.resize((size_t)3 * c.A);
        o.nodex.resize((size_t)24 * c.E);
        cac_hip_download_atoms(sl.h, c.A ? o.x.data() : nullptr, nullptr, nullptr);
        if (c.E) cac_hip_download_elements(sl.h, o.nodex.data(), nullptr, nullptr, nullptr);
      }
    } catch (const HipError& e) {
      o.r.status = CAC_HIP_EHIP;
    }
  }
  o.done = true;
  std::lock_guard<std::mutex> lk(s->mu);
  sl.busy = false;
  s->in_flight--;
  if (a.stepper >= 0) s->load_of[a.stepper]--;
  if (!msg.empty()) s->err = "replica " + std::to_string(a.req->id) + ": " + msg;
  s->out[a.req->id] = std::move(o);
  s->finished++;
  s->cv_done.notify_all();
  s->cv_work.notify_all();  // a slot is free for the loader
  // an idle stepping thread waits for work or for the last replica in flight
  // to finish once the loader is done: wake it, or it would sleep through
  // the server's shutdown
  s->cv_ready.notify_all();
}

// The loader thread: admits the head of the queue into a free slot when the
// budget allows and loads it there (upload, neighbor build, first force) on
// the slot's own stream, while the stepping thread keeps stepping the others.
void load_loop(cac_hip_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    std::unique_ptr<Request> q;
    int slot = -1;
    double held = 0.0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop && s->queue.empty()) {
          s->loader_done = true;
          s->cv_ready.notify_all();
          return;
        }
        if (!s->queue.empty()) {
          double h = 0.0;
          for (const Slot& sl : s->slots) h += sl.held;
          const double need = s->queue.front()->need;
          for (int k = 0; k < (int)s->slots.size(); k++) {
            const Slot& sl = s->slots[k];
            if (sl.busy) continue;
            // a replica alone on the device is always tried
            if (h - sl.held + std::max(sl.held, need) <= s->budget || s->in_flight == 0) {
              slot = k;
              break;
            }
          }
          if (slot >= 0) break;
        }
        s->cv_work.wait(lk);
      }
      q = std::move(s->queue.front());
      s->queue.pop_front();
      Slot& sl = s->slots[slot];
      sl.busy = true;
      sl.held = std::max(sl.held, q->need);
      for (const Slot& x : s->slots) held += x.held;
      s->in_flight++;
      s->peak_hbm = std::max(s->peak_hbm, held);
    }
    Active a;
    a.slot = slot;
    a.req = std::move(q);
    a.t_admit = s->now();
    std::string msg;
    const int rc = load(s->slots[slot], *a.req, msg);
    a.t_first = a.t_last = s->now();
    if (rc) {
      finish(s, a, rc, msg);
      continue;
    }
    mark_step(s, s->slots[slot]);  // the first step's interval starts when the load is done
    std::lock_guard<std::mutex> lk(s->mu);
    int best = 0;  // the stepping thread with the fewest replicas takes it
    for (int k = 1; k < (int)s->load_of.size(); k++)
      if (s->load_of[k] < s->load_of[best]) best = k;
    a.stepper = best;
    s->load_of[best]++;
    s->ready[best].push_back(std::move(a));
    s->cv_ready.notify_all();
  }
}

// The rebuilder thread: finishes the steps whose check asked for an outer
// rebuild (rebuild, pair passes, closing kick, as step_resolve does on a
// stepping thread), then returns each replica to its stepping thread.  A
// rebuild reads counts back to the host several times; done here, it does not
// hold up the other replicas of that stepping thread.
void rebuild_loop(cac_hip_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    Active a;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_rebuild.wait(lk, [&] { return !s->rebuild_q.empty() || s->rebuild_stop; });
      if (s->rebuild_q.empty()) return;
      a = std::move(s->rebuild_q.front());
      s->rebuild_q.pop_front();
    }
    Ctx& c = s->slots[a.slot].h->c;
    const int ev = a.done + 1 == a.req->r.nsteps ? 1 : 0;
    try {
      step_resolve(c, a.req->r.dt, ev, a.done + 1 < a.req->r.nsteps);
    } catch (const HipError& e) {
      finish(s, a, code_of(e), e.where);
      continue;
    }
    mark_step(s, s->slots[a.slot]);
    const double t = s->now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->interval_ms.push_back((float)(1000.0 * (t - a.t_last)));
    }
    a.t_last = t;
    a.done++;
    a.away = false;
    if (a.done >= a.req->r.nsteps) {
      finish(s, a, 0, "");
      continue;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const int me = a.stepper;
    s->ready[me].push_back(std::move(a));
    s->cv_ready.notify_all();
  }
}

// stream and event of a stepping thread's segmented pair passes
struct SegStream {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  ~SegStream() {
    if (done) (void)hipEventDestroy(done);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// The pair passes of the replicas in grp (each prepared: check resolved,
// lists rebuilt or re-pruned if asked) as segmented launches on the thread's
// own stream, then each replica's closing half-step on its own stream.  A
// replica's stream reaches the launch through its ev_seg event, and waits for
// the launch through seg.done.
void resolve_group(cac_hip_server* s, std::vector<Active*>& grp, SegStream& seg,
                   const std::function<void(Active&)>& stepped) {
  const int cap = max_segments();
  for (size_t b0 = 0; b0 < grp.size(); b0 += cap) {
    const size_t b1 = std::min(grp.size(), b0 + (size_t)cap);
    std::vector<Ctx*> cs;
    for (size_t i = b0; i < b1; i++) cs.push_back(&s->slots[grp[i]->slot].h->c);
    try {
      if (cs.size() == 1) {
        pair_passes(*cs[0], 0, 0, Gate());
      } else {
        for (Ctx* c : cs) {
          CAC_HIP(hipEventRecord(c->ev_seg, c->stream));
          CAC_HIP(hipStreamWaitEvent(seg.stream, c->ev_seg, 0));
        }
        pair_passes_segmented(cs.data(), (int)cs.size(), seg.stream);
        CAC_HIP(hipEventRecord(seg.done, seg.stream));
        for (Ctx* c : cs) CAC_HIP(hipStreamWaitEvent(c->stream, seg.done, 0));
        std::lock_guard<std::mutex> lk(s->mu);
        s->seg_launches++;
        s->seg_steps += (long)cs.size();
      }
    } catch (const HipError& e) {
      for (size_t i = b0; i < b1; i++) {
        grp[i]->done = -1;
        finish(s, *grp[i], code_of(e), e.where);
      }
      continue;
    }
    for (size_t i = b0; i < b1; i++) {
      Active& a = *grp[i];
      Ctx& c = s->slots[a.slot].h->c;
      try {
        step_close(c, a.req->r.dt, 0, a.done + 1 < a.req->r.nsteps);
      } catch (const HipError& e) {
        a.done = -1;
        finish(s, a, code_of(e), e.where);
        continue;
      }
      stepped(a);
    }
  }
}

// A stepping thread: takes the replicas the loader assigned to it, then runs
// rounds over them.  Several stepping threads share the host-side cost of
// queueing the launches; each owns its replicas' slots while it steps them.
// With CAC_HIP_SERVE_SEGMENT=1, replicas whose pair passes can run together
// (mixed-precision EAM, the default launch shape, not on an energy step, at
// most CAC_HIP_SERVE_SEGMENT_MAXP points) take them as segmented launches:
// one density, one embedding and one force launch for the group.
void serve(cac_hip_server* s, int me) {
  (void)hipSetDevice(s->device);
  std::vector<Active> act;
  std::vector<float> round_ms;
  // Off unless asked for: grouping makes every replica of the group wait for
  // the group's three launches, which costs more than the launches it saves
  // while the replicas' own streams already overlap on the device (512
  // replicas of 1.6k-198k atom-eq, 32 slots, saturated: 1.01e9 atom-eq
  // steps/s grouped against 1.41e9 alone; grouping only replicas of at most
  // 4,096 points: 1.37e9).
  bool segment = false;
  if (const char* e = std::getenv("CAC_HIP_SERVE_SEGMENT")) segment = std::atoi(e) != 0;
  long seg_maxp = 1L << 30;  // replicas of at most this many points join the groups
  if (const char* e = std::getenv("CAC_HIP_SERVE_SEGMENT_MAXP")) seg_maxp = std::atol(e);
  SegStream seg;
  if (segment) {
    if (hipStreamCreateWithFlags(&seg.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&seg.done, hipEventDisableTiming) != hipSuccess)
      segment = false;
  }
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (act.empty())
        s->cv_ready.wait(lk, [&] { return !s->ready[me].empty() || (s->loader_done && s->in_flight == 0); });
      for (Active& a : s->ready[me]) act.push_back(std::move(a));
      s->ready[me].clear();
      if (act.empty()) return;  // the loader has finished and nothing is in flight
      s->peak_active = std::max(s->peak_active, s->in_flight);
    }
    round_ms.clear();
    // one round: queue the first half of a step everywhere, then resolve
    for (Active& a : act) {
      Ctx& c = s->slots[a.slot].h->c;
      const int ev = a.done + 1 == a.req->r.nsteps ? 1 : 0;
      try {
        a.issued = step_issue(c, a.req->r.dt, ev);
      } catch (const HipError& e) {
        a.issued = false;
        a.done = -1;  // failed
        finish(s, a, code_of(e), e.where);
        continue;
      }
      if (!a.issued) {
        mark_step(s, s->slots[a.slot]);
        const double t = s->now();
        round_ms.push_back((float)(1000.0 * (t - a.t_last)));
        a.t_last = t;
        a.done++;
      }
    }
    auto stepped = [&](Active& a) {
      mark_step(s, s->slots[a.slot]);
      const double t = s->now();
      round_ms.push_back((float)(1000.0 * (t - a.t_last)));
      a.t_last = t;
      a.done++;
    };
    std::vector<Active*> grp;
    for (Active& a : act) {
      if (a.done < 0 || !a.issued) continue;
      Ctx& c = s->slots[a.slot].h->c;
      const int ev = a.done + 1 == a.req->r.nsteps ? 1 : 0;
      try {
        wait_check(c);
        if (c.h_steps[0] == (int)c.ntimestep) {  // an outer rebuild: the rebuilder takes the step
          a.away = true;
          continue;
        }
        if (segment && !ev && c.P <= seg_maxp && segmentable(c)) {
          step_prepare(c);
          grp.push_back(&a);
          continue;
        }
        step_resolve(c, a.req->r.dt, ev, a.done + 1 < a.req->r.nsteps);
      } catch (const HipError& e) {
        a.done = -1;
        finish(s, a, code_of(e), e.where);
        continue;
      }
      stepped(a);
    }
    if (!grp.empty()) resolve_group(s, grp, seg, stepped);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->interval_ms.insert(s->interval_ms.end(), round_ms.begin(), round_ms.end());
    }
    for (Active& a : act) {
      if (a.done < 0 || a.away) continue;
      if (a.done >= a.req->r.nsteps) {
        finish(s, a, 0, "");
        a.done = -1;
      }
    }
    {
      std::lock_guard<std::mutex> lk(s->mu);
      bool any = false;
      for (Active& a : act)
        if (a.away && a.done >= 0) {
          s->rebuild_q.push_back(std::move(a));
          a.done = -1;  // the moved-from entry leaves this thread's list
          any = true;
        }
      if (any) s->cv_rebuild.notify_all();
    }
    act.erase(std::remove_if(act.begin(), act.end(), [](const Active& a) { return a.done < 0; }), act.end());
  }
}

}  // namespace

extern "C" {

int cac_hip_server_create(cac_hip_server** out, int device, int slots, double hbm_budget_bytes) {
  if (!out) return CAC_HIP_EARG;
  *out = nullptr;
  if (slots < 1 || slots > 256) return CAC_HIP_EARG;
  auto* s = new cac_hip_server();
  s->device = device;
  s->t0 = Clock::now();
  for (int k = 0; k < slots; k++) {
    Slot sl;
    int rc = cac_hip_create(&sl.h, device, CAC_HIP_UNITS_METAL);
    for (int m = 0; m < MARKS && !rc; m++)
      // device-scope release: the mark times the step, not a system-scope write-back
      if (hipEventCreateWithFlags(&sl.marks[m], hipEventReleaseToDevice) != hipSuccess) rc = CAC_HIP_EHIP;
    if (rc) {
      s->slots.push_back(sl);
      for (Slot& x : s->slots) {
        for (hipEvent_t e : x.marks)
          if (e) (void)hipEventDestroy(e);
        if (x.h) cac_hip_destroy(x.h);
      }
      delete s;
      return rc;
    }
    s->slots.push_back(sl);
  }
  if (hbm_budget_bytes <= 0.0) {
    size_t fr = 0, tot = 0;
    (void)hipSetDevice(device);
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
    hbm_budget_bytes = 0.75 * (double)fr;
  }
  s->budget = hbm_budget_bytes;
  int nstep = 4;
  if (const char* e = std::getenv("CAC_HIP_SERVE_STEPPERS")) nstep = std::max(1, std::min(16, std::atoi(e)));
  nstep = std::min(nstep, slots);
  s->ready.resize(nstep);
  s->load_of.assign(nstep, 0);
  for (int k = 0; k < nstep; k++) s->steppers.emplace_back(serve, s, k);
  // one rebuilder keeps up at saturation (two measured no better: 1.41e9
  // against 1.42e9 atom-eq steps/s, p99 step interval 0.84-0.91 against
  // 0.87-1.07 ms)
  int nreb = 1;
  if (const char* e = std::getenv("CAC_HIP_SERVE_REBUILDERS")) nreb = std::max(1, std::min(8, std::atoi(e)));
  for (int k = 0; k < nreb; k++) s->rebuilders.emplace_back(rebuild_loop, s);
  s->loader = std::thread(load_loop, s);
  *out = s;
  return CAC_HIP_OK;
}

int cac_hip_server_destroy(cac_hip_server* s) {
  if (!s) return CAC_HIP_EARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_ready.notify_all();
  if (s->loader.joinable()) s->loader.join();
  for (std::thread& t : s->steppers)
    if (t.joinable()) t.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->rebuild_stop = true;  // the steppers have left, so no replica is away
  }
  s->cv_rebuild.notify_all();
  for (std::thread& t : s->rebuilders)
    if (t.joinable()) t.join();
  for (Slot& sl : s->slots) {
    cac_hip_destroy(sl.h);  // synchronizes the slot's stream first
    for (hipEvent_t e : sl.marks)
      if (e) (void)hipEventDestroy(e);
  }
  delete s;
  return CAC_HIP_OK;
}

// The service threads may replace the message at any time, so the caller gets
// a copy taken under the lock, held per calling thread: the pointer stays
// valid until that thread calls again.
const char* cac_hip_server_last_error(const cac_hip_server* s) {
  if (!s) return "null server";
  thread_local std::string copy;
  std::lock_guard<std::mutex> lk(s->mu);
  copy = s->err;
  return copy.c_str();
}

int cac_hip_server_pair_eam(cac_hip_server* s, int nrho, double drho, int nr, double dr, double cut,
                            const double* frho, const double* rhor, const double* z2r) {
  if (!s) return CAC_HIP_EARG;
  if (s->submitted) return set_err(s, CAC_HIP_ESTATE, "set the pair style before the first submit");
  for (Slot& sl : s->slots) {
    const int rc = cac_hip_pair_eam(sl.h, nrho, drho, nr, dr, cut, frho, rhor, z2r);
    if (rc) return set_err(s, rc, cac_hip_last_error(sl.h));
  }
  s->pair_set = true;
  return CAC_HIP_OK;
}

int cac_hip_server_pair_lj(cac_hip_server* s, double eps, double sigma, double cut) {
  if (!s) return CAC_HIP_EARG;
  if (s->submitted) return set_err(s, CAC_HIP_ESTATE, "set the pair style before the first submit");
  for (Slot& sl : s->slots) {
    const int rc = cac_hip_pair_lj(sl.h, eps, sigma, cut);
    if (rc) return set_err(s, rc, cac_hip_last_error(sl.h));
  }
  s->pair_set = true;
  return CAC_HIP_OK;
}

int cac_hip_server_submit(cac_hip_server* s, const cac_hip_request* req, long* id) {
  if (!s || !req) return CAC_HIP_EARG;
  if (!s->pair_set) return set_err(s, CAC_HIP_ESTATE, "no pair style");
  const cac_hip_request& r = *req;
  if (r.ntypes < 1 || !r.mass || r.natoms < 0 || r.nelements < 0 || r.natoms + r.nelements == 0)
    return set_err(s, CAC_HIP_EARG, "bad replica sizes");
  if (r.nsteps < 1) return set_err(s, CAC_HIP_EARG, "a replica runs at least one step");
  if (r.natoms > 0 && (!r.x || !r.type)) return set_err(s, CAC_HIP_EARG, "bad atom arrays");
  if (r.nelements > 0 && (!r.ctype || !r.nodex)) return set_err(s, CAC_HIP_EARG, "bad element arrays");
  auto q = std::make_unique<Request>();
  q->r = r;
  q->mass.assign(r.mass, r.mass + r.ntypes);
  const size_t na = 3 * (size_t)r.natoms, ne = 24 * (size_t)r.nelements;
  if (r.natoms) {
    q->x.assign(r.x, r.x + na);
    q->v.assign(na, 0.0);
    if (r.v) q->v.assign(r.v, r.v + na);
    q->type.assign(r.type, r.type + r.natoms);
  }
  if (r.nelements) {
    q->nodex.assign(r.nodex, r.nodex + ne);
    q->nodev.assign(ne, 0.0);
    if (r.nodev) q->nodev.assign(r.nodev, r.nodev + ne);
    q->ctype.assign(r.ctype, r.ctype + r.nelements);
  }
  q->etype.assign(r.nelements, 1);
  q->r.mass = q->mass.data();
  q->r.x = q->x.data();
  q->r.v = q->v.data();
  q->r.type = q->type.data();
  q->r.nodex = q->nodex.data();
  q->r.nodev = q->nodev.data();
  q->r.ctype = q->ctype.data();
  q->need = need_bytes(r);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    q->id = s->next_id++;
    q->t_submit = s->now();
    if (id) *id = q->id;
    s->submitted = true;
    s->queue.push_back(std::move(q));
  }
  s->cv_work.notify_all();
  return CAC_HIP_OK;
}

int cac_hip_server_drain(cac_hip_server* s) {
  if (!s) return CAC_HIP_EARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->finished == s->next_id; });
  return CAC_HIP_OK;
}

int cac_hip_server_result(cac_hip_server* s, long id, cac_hip_result* res) {
  if (!s || !res) return CAC_HIP_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->out.find(id);
  if (it == s->out.end() || !it->second.done) {
    s->err = "no finished replica with that id";
    return CAC_HIP_ESTATE;
  }
  *res = it->second.r;
  return CAC_HIP_OK;
}

int cac_hip_server_state(cac_hip_server* s, long id, double* x, double* nodex) {
  if (!s) return CAC_HIP_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->out.find(id);
  if (it == s->out.end() || !it->second.done) {
    s->err = "no finished replica with that id";
    return CAC_HIP_ESTATE;
  }
  const Outcome& o = it->second;
  if (o.x.empty() && o.nodex.empty()) {
    s->err = "replica was not submitted with keep_state";
    return CAC_HIP_ESTATE;
  }
  if (x && !o.x.empty()) std::memcpy(x, o.x.data(), sizeof(double) * o.x.size());
  if (nodex && !o.nodex.empty()) std::memcpy(nodex, o.nodex.data(), sizeof(double) * o.nodex.size());
  return CAC_HIP_OK;
}

int cac_hip_server_segments(cac_hip_server* s, long* launches, long* replica_steps) {
  if (!s) return CAC_HIP_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (launches) *launches = s->seg_launches;
  if (replica_steps) *replica_steps = s->seg_steps;
  return CAC_HIP_OK;
}

int cac_hip_server_device_intervals(cac_hip_server* s, double* p50_ms, double* p99_ms, long* n) {
  if (!s) return CAC_HIP_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<float> v = s->device_ms;
  auto pct = [&](double q) -> double {
    if (v.empty()) return 0.0;
    const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + (long)k, v.end());
    return v[k];
  };
  if (p50_ms) *p50_ms = pct(0.50);
  if (p99_ms) *p99_ms = pct(0.99);
  if (n) *n = (long)s->device_ms.size();
  return CAC_HIP_OK;
}

int cac_hip_server_stats(cac_hip_server* s, double* p50_ms, double* p99_ms, long* nsteps, double* atom_steps_per_s,
                         int* peak_active, double* peak_hbm_bytes) {
  if (!s) return CAC_HIP_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<float> v = s->interval_ms;
  auto pct = [&](double q) -> double {
    if (v.empty()) return 0.0;
    const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + (long)k, v.end());
    return v[k];
  };
  if (p50_ms) *p50_ms = pct(0.50);
  if (p99_ms) *p99_ms = pct(0.99);
  if (nsteps) *nsteps = (long)s->interval_ms.size();
  // aggregate rate: atom-eq steps of the finished replicas over the union of
  // their [admission, done] intervals
  std::vector<std::pair<double, double>> span;
  double work = 0.0;
  for (const auto& kv : s->out) {
    const cac_hip_result& r = kv.second.r;
    if (!kv.second.done || r.status) continue;
    span.push_back({r.t_admit, r.t_done});
    work += (double)r.nsteps * (double)r.atom_equivalents;
  }
  std::sort(span.begin(), span.end());
  double busy = 0.0, lo = -1.0, hi = -1.0;
  for (const auto& iv : span) {
    if (iv.first > hi) {
      busy += hi - lo;
      lo = iv.first;
      hi = iv.second;
    } else {
      hi = std::max(hi, iv.second);
    }
  }
  busy += hi - lo;
  if (atom_steps_per_s) *atom_steps_per_s = busy > 0.0 ? work / busy : 0.0;
  if (peak_active) *peak_active = s->peak_active;
  if (peak_hbm_bytes) *peak_hbm_bytes = s->peak_hbm;
  return CAC_HIP_OK;
}

}  // extern "C"
