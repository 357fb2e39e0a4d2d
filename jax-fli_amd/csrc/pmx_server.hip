// Forward-simulation request server (pmx_server_* in include/pmx.h): the reference runs one jitted
// simulation per call (run_simulations, src/jax_fli/scripts/fli_simulate.py:512-585; each one
// lpt + nbody, pm/nbody.py:28-171); this keeps many in flight on one GPU.
//
// * Size classes. One class per mesh size n, with B slots. A slot's particles, meshes, spectra and
//   sort buffers live in one arena of a batched plan (pmx_plan_create_batched), so a request stays
//   resident in HBM from its initial conditions to its last step. The HBM budget fixes B per class;
//   a request is admitted (per class, in arrival order) when a slot is free.
// * Batched steps. Every fused PM step phase (sorted paint, the potential's FFT passes, gradient
//   stencil, readout + kick + drift) is one launch over all slots of a class; the kernels read each
//   slot's coefficients of its current step from a device table and skip idle slots.
// * Graphs. One step of a class (coefficient fetch, the phases, step counter update) is captured once
//   as a HIP graph; the scheduler replays it while the class has live slots. No memset is captured.
// * Threads. Host preparation of a request (EH98 P(k) with its sigma8 integral, growth table, time
//   grid and KDK coefficients; pmx_cosmo.h) runs on a pool of host threads; one scheduler thread owns
//   every device call: admissions (white noise, colouring and LPT of the new slot), step replays
//   (a few ahead of the device per class), and each finished request's density paint and copy-out.
// Kernels are deterministic and slots do not interact, so a result is the same bit for bit whatever
// else is in flight.
#include "pmx_internal.h"
#include "pmx_cosmo.h"
#include "../../include/pmx.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int kMaxOps = 256;  // steps per request (KDK, dt0 >= 1/250 from a = 0 to 1)
constexpr int kTab = 128;     // colouring table length
constexpr int kAhead = 3;     // step replays queued ahead of the device per class

// the caller region of a slot (after the plan's own buffers): particles, white noise, initial field,
// colouring table, per-step coefficients
struct SlotRegion {
  size_t x, v, noise, ic, ktab, ptab, coef, bytes;
};
SlotRegion slot_region(int n) {
  const size_t ncell = (size_t)n * n * n;
  SlotRegion r;
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t at = off;
    off += (b + 255) & ~(size_t)255;
    return at;
  };
  r.x = take(3 * ncell * 4);
  r.v = take(3 * ncell * 4);
  r.noise = take(ncell * 4);
  r.ic = take(ncell * 4);
  r.ktab = take(kTab * 8);
  r.ptab = take(kTab * 8);
  r.coef = take((size_t)kMaxOps * 4 * 4);
  r.bytes = off;
  return r;
}

__global__ void k_slot_set(int* live, int* tick, int* nops, int z, int on, int n_ops) {
  live[z] = on;
  if (on) {
    tick[z] = 0;
    nops[z] = n_ops;
  }
}

// per slot: runs this step (live, steps left) and its coefficients {unused, kick, drift, potential scale}
__global__ void k_tick_prep(const int* live, const int* tick, const int* nops, const char* coef0, size_t stride,
                            float* cur, int* live_now, int B) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= B) return;
  const int run = live[z] && tick[z] < nops[z];
  live_now[z] = run;
  if (run) {
    const float* c = reinterpret_cast<const float*>(coef0 + (size_t)z * stride) + 4 * tick[z];
    for (int k = 0; k < 4; ++k) cur[4 * z + k] = c[k];
  }
}

__global__ void k_tick_advance(const int* live_now, int* tick, int B) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z < B) tick[z] += live_now[z];
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct Request {
  uint64_t id = 0;
  pmx_request_t r{};
  float* out_x = nullptr;
  float* out_rho = nullptr;
  int cls = -1, slot = -1;
  std::atomic<int> state{PMX_REQ_QUEUED};
  bool prepared = false;  // host tables ready (guarded by the server mutex)
  int prep_rc = 0;
  pmxc::RequestPlan plan;
  double t_submit = 0, t_admit = 0, t_done = 0;
  hipEvent_t done = nullptr;
};

struct Class {
  int n = 0, B = 0;
  pmx_plan_t* bp = nullptr;
  size_t extra_off = 0;
  SlotRegion reg{};
  std::vector<pmx_plan_t> views;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int *live = nullptr, *tick = nullptr, *nops = nullptr, *live_now = nullptr;
  float* cur = nullptr;
  struct Slot {
    Request* req = nullptr;
    int ticks = 0, nops = 0;
  };
  std::vector<Slot> slots;
  std::vector<char*> staging;       // pinned, per slot: ktab, ptab, kphys x/y/z, coefficients
  std::deque<Request*> fifo;        // this size's requests in arrival order (guarded by the mutex)
  std::deque<hipEvent_t> inflight;  // after each queued step replay
  std::vector<hipEvent_t> free_events;
  std::vector<Request*> finishing;  // copy-out queued, completion not yet seen

  char* slot_base(int z) const { return reinterpret_cast<char*>(bp->arena) + (size_t)z * bp->slots.stride + extra_off; }
  size_t staging_bytes() const { return (size_t)(2 * kTab + 2 * n + n / 2 + 1) * 8 + (size_t)kMaxOps * 16; }
};

}  // namespace

struct pmx_server_t {
  int device = 0;
  std::vector<Class> classes;
  std::mutex mu;
  std::condition_variable cv_prep, cv_done;
  std::deque<Request*> prep_queue;
  std::map<uint64_t, Request*> reqs;
  uint64_t next_id = 1;
  std::vector<std::thread> preppers;
  std::thread worker;
  std::atomic<bool> stop{false};
  std::atomic<int> running{0};  // requests holding a slot
  int peak_running = 0;
  std::string error;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  size_t device_bytes = 0;
};

namespace {

int class_of(const pmx_server_t* s, int mesh) {
  for (size_t i = 0; i < s->classes.size(); ++i)
    if (s->classes[i].n == mesh) return (int)i;
  return -1;
}

hipEvent_t take_event(Class& c) {
  if (!c.free_events.empty()) {
    hipEvent_t e = c.free_events.back();
    c.free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
  return e;
}

#define SRV_TRY(expr)           \
  do {                          \
    int rc_ = (expr);           \
    if (rc_) return rc_;        \
  } while (0)
#define SRV_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return pmx_set_error(3, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Capture one step of the class: coefficients of each live slot's current step, the batched phases,
// the step counters.
int capture_step(Class& c) {
  pmx_plan_t* bp = c.bp;
  bp->slots.live = c.live_now;
  bp->slots.coef = c.cur;
  PartArgs a{};
  a.x = reinterpret_cast<float*>(c.slot_base(0) + c.reg.x);
  a.v = nullptr;
  a.npx = a.npy = a.npz = c.n;
  a.mode = 0;
  a.sx = a.sy = a.sz = 1.f;
  const int g = (c.B + 63) / 64;
  SRV_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
  hipLaunchKernelGGL(k_tick_prep, dim3(g), dim3(64), 0, c.stream, c.live, c.tick, c.nops,
                     (const char*)(c.slot_base(0) + c.reg.coef), bp->slots.stride, c.cur, c.live_now, c.B);
  int rc = pmx_check_launch("tick_prep");
  if (!rc) rc = pmx_launch_paint(bp, a, nullptr, 1.f, 0.f, bp->rho, 0, 0, c.stream);
  if (!rc) rc = pmx_pot_force_mesh(bp, bp->rho, 1.f, c.stream);
  if (!rc) {
    a.v = reinterpret_cast<float*>(c.slot_base(0) + c.reg.v);
    rc = pmx_launch_readout(bp, a, bp->fmesh, 3, nullptr, 0.f, 0.f, 2 | 4, c.stream);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_tick_advance, dim3(g), dim3(64), 0, c.stream, c.live_now, c.tick, c.B);
    rc = pmx_check_launch("tick_advance");
  }
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c.stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return pmx_set_error(3, std::string("step capture: ") + hipGetErrorString(e));
  c.graph = graph;
  SRV_HIP(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
  return 0;
}

int create_class(Class& c, int n, int B) {
  c.n = n;
  c.B = B;
  c.reg = slot_region(n);
  SRV_TRY(pmx_plan_create_batched(&c.bp, n, B, c.reg.bytes, &c.extra_off));
  for (int z = 0; z < B; ++z) {
    c.views.push_back(pmx_slot_view(c.bp, z));
    c.views.back().box_set = true;  // each admission uploads its request's k tables first
  }
  c.slots.assign(B, Class::Slot{});
  SRV_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  SRV_HIP(hipMalloc((void**)&c.live, 4 * sizeof(int) * (size_t)B));
  c.tick = c.live + B;
  c.nops = c.live + 2 * B;
  c.live_now = c.live + 3 * B;
  SRV_HIP(hipMalloc((void**)&c.cur, 4 * sizeof(float) * (size_t)B));
  SRV_HIP(hipMemset(c.live, 0, 4 * sizeof(int) * (size_t)B));
  SRV_HIP(hipMemset(c.cur, 0, 4 * sizeof(float) * (size_t)B));
  for (int z = 0; z < B; ++z) {
    char* h = nullptr;
    SRV_HIP(hipHostMalloc((void**)&h, c.staging_bytes(), hipHostMallocDefault));
    c.staging.push_back(h);
  }
  return capture_step(c);
}

void destroy_class(Class& c) {
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  if (c.exec) (void)hipGraphExecDestroy(c.exec);
  if (c.graph) (void)hipGraphDestroy(c.graph);
  for (hipEvent_t e : c.inflight) (void)hipEventDestroy(e);
  for (hipEvent_t e : c.free_events) (void)hipEventDestroy(e);
  for (char* h : c.staging) (void)hipHostFree(h);
  if (c.live) (void)hipFree(c.live);
  if (c.cur) (void)hipFree(c.cur);
  if (c.bp) pmx_plan_destroy(c.bp);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  c = Class{};
}

// the estimator wavevectors of pmx_plan_set_box for a cubic box
void kphys_table(int n, int count, double L, double* t) {
  const double val = 1.0 / (double)n;
  for (int i = 0; i < count; ++i) {
    const int f = (i < (n + 1) / 2) ? i : i - n;
    t[i] = (2.0 * M_PI * (double)n / L) * ((double)f * val);
  }
}

// Initial conditions of a newly admitted request in slot z: tables in, white noise, colouring, LPT,
// then the slot goes live for the step replays.
int admit(Class& c, int z, Request* q) {
  const int n = c.n;
  char* h = c.staging[z];
  double* ktab = reinterpret_cast<double*>(h);
  double* ptab = ktab + kTab;
  double* kx = ptab + kTab;
  double* ky = kx + n;
  double* kz = ky + n;
  float* coef = reinterpret_cast<float*>(kz + n / 2 + 1);
  const pmxc::RequestPlan& P = q->plan;
  std::copy(P.ktab.begin(), P.ktab.end(), ktab);
  std::copy(P.ptab.begin(), P.ptab.end(), ptab);
  kphys_table(n, n, q->r.box, kx);
  kphys_table(n, n, q->r.box, ky);
  kphys_table(n, n / 2 + 1, q->r.box, kz);
  const size_t ncell = (size_t)n * n * n;
  // the potential pass's scale, as pot_only forms it from the force scale
  const float pot = (float)(P.force_scale / (12.0 * (double)ncell));
  const int nops = (int)P.ops.size();
  for (int i = 0; i < nops; ++i) {
    coef[4 * i] = 0.f;
    coef[4 * i + 1] = P.ops[i].kick;
    coef[4 * i + 2] = std::isnan(P.ops[i].post) ? 0.f : P.ops[i].post;
    coef[4 * i + 3] = pot;
  }
  pmx_plan_t* v = &c.views[z];
  char* base = c.slot_base(z);
  SRV_HIP(hipMemcpyAsync(base + c.reg.ktab, ktab, 2 * kTab * 8, hipMemcpyHostToDevice, c.stream));
  SRV_HIP(hipMemcpyAsync(v->kphys_x, kx, (size_t)n * 8, hipMemcpyHostToDevice, c.stream));
  SRV_HIP(hipMemcpyAsync(v->kphys_y, ky, (size_t)n * 8, hipMemcpyHostToDevice, c.stream));
  SRV_HIP(hipMemcpyAsync(v->kphys_z, kz, (size_t)(n / 2 + 1) * 8, hipMemcpyHostToDevice, c.stream));
  SRV_HIP(hipMemcpyAsync(base + c.reg.coef, coef, (size_t)nops * 16, hipMemcpyHostToDevice, c.stream));
  float* noise = reinterpret_cast<float*>(base + c.reg.noise);
  float* ic = reinterpret_cast<float*>(base + c.reg.ic);
  float* x = reinterpret_cast<float*>(base + c.reg.x);
  float* vel = reinterpret_cast<float*>(base + c.reg.v);
  SRV_TRY(pmx_white_noise(v, q->r.seed, noise, c.stream));
  SRV_TRY(pmx_color_field(v, noise, reinterpret_cast<double*>(base + c.reg.ktab),
                          reinterpret_cast<double*>(base + c.reg.ptab), kTab, ic, c.stream));
  SRV_TRY(pmx_lpt1(v, ic, P.lpt_growth, P.lpt_vel, x, vel, c.stream));
  hipLaunchKernelGGL(k_slot_set, dim3(1), dim3(1), 0, c.stream, c.live, c.tick, c.nops, z, 1, nops);
  SRV_TRY(pmx_check_launch("slot_set"));
  c.slots[z].req = q;
  c.slots[z].ticks = 0;
  c.slots[z].nops = nops;
  return 0;
}

// A slot whose steps are all queued: idle it, paint the final density into the caller's mesh, copy
// the displacements out, mark completion. The slot is free for the next admission at once (stream
// order keeps the copy before any later use).
int finish(Class& c, int z) {
  Request* q = c.slots[z].req;
  pmx_plan_t* v = &c.views[z];
  char* base = c.slot_base(z);
  hipLaunchKernelGGL(k_slot_set, dim3(1), dim3(1), 0, c.stream, c.live, c.tick, c.nops, z, 0, 0);
  SRV_TRY(pmx_check_launch("slot_set"));
  const float* x = reinterpret_cast<const float*>(base + c.reg.x);
  SRV_TRY(pmx_paint_cic(v, x, c.n, c.n, c.n, 0, nullptr, 1.f, q->out_rho, 0, c.stream));
  SRV_HIP(hipMemcpyAsync(q->out_x, x, (size_t)c.n * c.n * c.n * 12, hipMemcpyDeviceToDevice, c.stream));
  q->done = take_event(c);
  SRV_HIP(hipEventRecord(q->done, c.stream));
  c.finishing.push_back(q);
  c.slots[z].req = nullptr;
  return 0;
}

void fail_all(pmx_server_t* s, const std::string& msg) {
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->error.empty()) s->error = msg;
  for (auto& kv : s->reqs) {
    const int st = kv.second->state.load();
    if (st != PMX_REQ_DONE) kv.second->state = PMX_REQ_FAILED;
  }
  s->cv_done.notify_all();
}

void prep_loop(pmx_server_t* s) {
  for (;;) {
    Request* q = nullptr;
    {
      std::unique_lock<std::mutex> lock(s->mu);
      s->cv_prep.wait(lock, [&] { return s->stop.load() || !s->prep_queue.empty(); });
      if (s->stop.load()) return;
      q = s->prep_queue.front();
      s->prep_queue.pop_front();
    }
    const pmx_request_t& r = q->r;
    const pmxc::Params p{r.Omega_c, r.Omega_b, r.h, r.n_s, r.sigma8, 0.0, -1.0, 0.0};
    pmxc::RequestPlan plan = pmxc::plan_request(p, r.mesh, r.box, r.a0, r.a1, r.dt0);
    std::lock_guard<std::mutex> lock(s->mu);
    q->plan = std::move(plan);
    q->prep_rc = (int)q->plan.ops.size() > kMaxOps ? 1 : 0;
    q->prepared = true;
  }
}

void worker_loop(pmx_server_t* s) {
  if (hipSetDevice(s->device) != hipSuccess) {
    fail_all(s, "server thread: hipSetDevice failed");
    return;
  }
  for (;;) {
    bool did = false;
    for (Class& c : s->classes) {
      // completions
      for (size_t i = 0; i < c.finishing.size();) {
        Request* q = c.finishing[i];
        if (hipEventQuery(q->done) == hipSuccess) {
          {
            std::lock_guard<std::mutex> lock(s->mu);
            q->t_done = now_s(s->t0);
            q->state = PMX_REQ_DONE;
          }
          s->cv_done.notify_all();
          c.free_events.push_back(q->done);
          q->done = nullptr;
          c.finishing.erase(c.finishing.begin() + (long)i);
          did = true;
        } else {
          ++i;
        }
      }
      while (!c.inflight.empty() && hipEventQuery(c.inflight.front()) == hipSuccess) {
        c.free_events.push_back(c.inflight.front());
        c.inflight.pop_front();
      }
      // slots whose every step is queued
      for (int z = 0; z < c.B; ++z) {
        if (c.slots[z].req && c.slots[z].ticks == c.slots[z].nops) {
          const int rc = finish(c, z);
          if (rc) return fail_all(s, std::string("finishing a request: ") + pmx_last_error());
          s->running--;  // the slot is free for the next admission
          did = true;
        }
      }
      // admissions, this size's arrival order
      for (int z = 0; z < c.B && !s->stop.load(); ++z) {
        if (c.slots[z].req) continue;
        Request* q = nullptr;
        {
          std::lock_guard<std::mutex> lock(s->mu);
          if (c.fifo.empty() || !c.fifo.front()->prepared) break;
          q = c.fifo.front();
          c.fifo.pop_front();
          if (q->prep_rc) {
            q->state = PMX_REQ_FAILED;
            s->cv_done.notify_all();
            continue;
          }
          q->t_admit = now_s(s->t0);
          q->state = PMX_REQ_RUNNING;
          q->slot = z;
          s->running++;
          s->peak_running = std::max(s->peak_running, s->running.load());
        }
        const int rc = admit(c, z, q);
        if (rc) return fail_all(s, std::string("admitting a request: ") + pmx_last_error());
        did = true;
      }
      // one step of every live slot
      bool any = false;
      for (int z = 0; z < c.B; ++z) any = any || (c.slots[z].req && c.slots[z].ticks < c.slots[z].nops);
      if (any && (int)c.inflight.size() < kAhead) {
        if (hipGraphLaunch(c.exec, c.stream) != hipSuccess) return fail_all(s, "step replay failed");
        hipEvent_t e = take_event(c);
        (void)hipEventRecord(e, c.stream);
        c.inflight.push_back(e);
        for (int z = 0; z < c.B; ++z)
          if (c.slots[z].req && c.slots[z].ticks < c.slots[z].nops) c.slots[z].ticks++;
        did = true;
      }
    }
    if (!did) {
      bool busy = false;
      for (Class& c : s->classes) {
        busy = busy || !c.inflight.empty() || !c.finishing.empty();
        for (const Class::Slot& sl : c.slots) busy = busy || sl.req != nullptr;
      }
      if (s->stop.load() && !busy) return;
      std::unique_lock<std::mutex> lock(s->mu);
      bool waiting = false;
      for (Class& c : s->classes) waiting = waiting || !c.fifo.empty();
      if (!waiting && !busy && !s->stop.load()) {
        s->cv_done.wait_for(lock, std::chrono::milliseconds(1));  // woken by submissions
      } else {
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
  }
}

}  // namespace

extern "C" {

int pmx_server_create(pmx_server* out, const int* meshes, const int* slots, int nclass, size_t hbm_budget,
                      int prep_threads) {
  if (!out || !meshes || nclass < 1) return pmx_set_error(1, "bad server arguments");
  *out = nullptr;
  for (int i = 0; i < nclass; ++i) {
    if (!pmx_fft_len_ok(meshes[i])) return pmx_set_error(2, "unsupported mesh size " + std::to_string(meshes[i]));
    for (int j = 0; j < i; ++j)
      if (meshes[j] == meshes[i]) return pmx_set_error(1, "mesh sizes must be distinct");
  }
  pmx_server_t* s = new pmx_server_t();
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return pmx_set_error(3, "hipGetDevice failed");
  }
  s->classes.resize(nclass);
  for (int i = 0; i < nclass; ++i) {
    const int n = meshes[i];
    const size_t per_slot = pmx_batched_slot_bytes(n, slot_region(n).bytes);
    int B = slots ? slots[i] : 0;
    if (B <= 0) {
      const size_t share = hbm_budget / (size_t)nclass;
      B = (int)std::min<size_t>(64, std::max<size_t>(1, share / per_slot));
    }
    const int rc = create_class(s->classes[i], n, B);
    if (rc) {
      for (Class& c : s->classes) destroy_class(c);
      delete s;
      return rc;
    }
    s->device_bytes += s->classes[i].bp->workspace_bytes;
  }
  const int np = std::max(1, std::min(prep_threads > 0 ? prep_threads : 4, 64));
  for (int i = 0; i < np; ++i) s->preppers.emplace_back(prep_loop, s);
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return 0;
}

int pmx_server_class(pmx_server s, int mesh, int* slots, size_t* slot_bytes) {
  if (!s) return pmx_set_error(1, "null server");
  const int c = class_of(s, mesh);
  if (c < 0) return pmx_set_error(1, "no class for mesh size " + std::to_string(mesh));
  if (slots) *slots = s->classes[c].B;
  if (slot_bytes) *slot_bytes = s->classes[c].bp->slots.stride;
  return 0;
}

int pmx_server_submit(pmx_server s, const pmx_request_t* req, float* displacement, float* density, uint64_t* id) {
  if (!s || !req || !displacement || !density || !id) return pmx_set_error(1, "null argument");
  const int c = class_of(s, req->mesh);
  if (c < 0) return pmx_set_error(1, "no class for mesh size " + std::to_string(req->mesh));
  if (!(req->box > 0) || !(req->a0 > 0) || !(req->a1 > req->a0) || !(req->dt0 > 0) || !(req->Omega_c > 0) ||
      !(req->sigma8 > 0) || !(req->h > 0))
    return pmx_set_error(1, "bad request parameters");
  Request* q = new Request();
  q->r = *req;
  q->out_x = displacement;
  q->out_rho = density;
  q->cls = c;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (!s->error.empty()) {
      delete q;
      return pmx_set_error(3, "server failed: " + s->error);
    }
    q->id = s->next_id++;
    q->t_submit = now_s(s->t0);
    s->reqs[q->id] = q;
    s->classes[c].fifo.push_back(q);
    s->prep_queue.push_back(q);
    *id = q->id;
  }
  s->cv_prep.notify_one();
  s->cv_done.notify_all();  // wakes an idle scheduler
  return 0;
}

int pmx_server_poll(pmx_server s, uint64_t id, int* state) {
  if (!s || !state) return pmx_set_error(1, "null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return pmx_set_error(1, "unknown request id");
  *state = it->second->state.load();
  return 0;
}

int pmx_server_wait(pmx_server s, uint64_t id, double timeout_s) {
  if (!s) return pmx_set_error(1, "null server");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return pmx_set_error(1, "unknown request id");
  Request* q = it->second;
  auto ready = [&] { return q->state.load() == PMX_REQ_DONE || q->state.load() == PMX_REQ_FAILED; };
  if (timeout_s > 0) {
    if (!s->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), ready))
      return pmx_set_error(5, "request not finished within the timeout");
  } else {
    s->cv_done.wait(lock, ready);
  }
  if (q->state.load() == PMX_REQ_FAILED)
    return pmx_set_error(3, "request failed: " + (s->error.empty() ? std::string("too many steps") : s->error));
  return 0;
}

int pmx_server_request_times(pmx_server s, uint64_t id, double times[3]) {
  if (!s || !times) return pmx_set_error(1, "null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return pmx_set_error(1, "unknown request id");
  times[0] = it->second->t_submit;
  times[1] = it->second->t_admit;
  times[2] = it->second->t_done;
  return 0;
}

int pmx_server_stats(pmx_server s, uint64_t since_id, pmx_server_stats_t* out) {
  if (!s || !out) return pmx_set_error(1, "null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  std::memset(out, 0, sizeof(*out));
  std::vector<double> lat;
  double first = 1e300, last = 0.0;
  for (auto it = s->reqs.lower_bound(since_id); it != s->reqs.end(); ++it) {
    const Request* q = it->second;
    out->submitted++;
    const int st = q->state.load();
    if (st == PMX_REQ_DONE) {
      out->completed++;
      lat.push_back((q->t_done - q->t_submit) * 1e3);
      last = std::max(last, q->t_done);
    } else if (st == PMX_REQ_FAILED) {
      out->failed++;
    } else if (st == PMX_REQ_RUNNING) {
      out->running++;
    } else {
      out->queued++;
    }
    first = std::min(first, q->t_submit);
  }
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double f) {  // linear interpolation between order statistics
      const double pos = f * (double)(lat.size() - 1);
      const size_t i = (size_t)pos;
      const double t = pos - (double)i;
      return i + 1 < lat.size() ? lat[i] * (1 - t) + lat[i + 1] * t : lat[i];
    };
    out->p50_ms = pct(0.5);
    out->p99_ms = pct(0.99);
    double sum = 0;
    for (double v : lat) sum += v;
    out->mean_ms = sum / (double)lat.size();
    out->max_ms = lat.back();
    out->wall_s = last - first;
    out->throughput_req_s = out->wall_s > 0 ? (double)out->completed / out->wall_s : 0.0;
  }
  out->peak_running = s->peak_running;
  out->device_bytes = s->device_bytes;
  return 0;
}

int pmx_server_destroy(pmx_server s) {
  if (!s) return 0;
  s->stop = true;
  s->cv_prep.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->preppers) t.join();
  if (s->worker.joinable()) s->worker.join();
  for (Class& c : s->classes) destroy_class(c);
  for (auto& kv : s->reqs) delete kv.second;
  delete s;
  return 0;
}

}  // extern "C"
