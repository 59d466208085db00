// Serving engine: continuous batching of subset-eigenvalue requests against a
// device-resident G. No reference counterpart (the reference resets the
// device and re-uploads G on every call, ParaLanczos.cu:1837-1858).
//
// A request is a batch of subsets whose index sets, sizes and output buffers
// live in device memory (the caller's). Requests queue in arrival order. A
// scheduler thread packs the queue's head into a step while it fits the
// step's HBM arena (subsets, indices, request-table entries) and its work
// budget (LDS bytes the subsets' kernels occupy, times their Lanczos steps),
// then on the step's own stream: copies the step's request table to the
// device, assembles the requests into one CSR batch in the arena (one
// kernel), runs the tuned binned solve over it, and scatters each subset's
// results to its request's outputs (one kernel). `depth` steps are in flight
// at once, each on its own stream, so requests that arrive while others run
// are packed into the next step and their work overlaps the running steps on
// the device. A completion thread retires steps as their events complete and
// stamps each request's latency. A subset's arithmetic depends only on its
// size k (the variant is fixed per k), so its results are the same whether it
// runs alone or in a mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/paralanczos.h"
#include "pl_common.hpp"
#include "pl_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

// request-table entry of a step (device copy of the host table)
struct ReqEntry {
  const int32_t *idx;
  const int64_t *off;  // request-local CSR offsets (off[0] = 0), or nullptr
  int64_t nsub, nidx;
  int64_t sub_base, idx_base;  // placement in the step's arena
  int32_t kuni;                // subset size when off == nullptr
  int32_t unused;
  float *lmin, *lmax, *delta;
  uint32_t *status;
};

// last entry r with base(r) <= x (bases ascending, base(0) = 0)
__device__ __forceinline__ int find_entry(const ReqEntry *t, int nreq, int64_t x, bool by_sub) {
  int lo = 0, hi = nreq - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t b = by_sub ? t[mid].sub_base : t[mid].idx_base;
    if (b <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// the step's CSR batch: concatenated index sets and rebased offsets
__global__ void assemble_kernel(const ReqEntry *t, int nreq, int64_t tot_sub, int64_t tot_idx,
                                int32_t *idx, int64_t *off) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t e = g; e < tot_idx; e += stride) {
    const ReqEntry &r = t[find_entry(t, nreq, e, false)];
    idx[e] = r.idx[e - r.idx_base];
  }
  for (int64_t s = g; s <= tot_sub; s += stride) {
    if (s == tot_sub) {
      off[s] = tot_idx;
      continue;
    }
    const ReqEntry &r = t[find_entry(t, nreq, s, true)];
    const int64_t l = s - r.sub_base;
    // clamped into the request's own index range, and the first subset of a
    // request starts at its own base whatever off[0] holds: a request's
    // subsets end where the next request's first one starts, so offsets that
    // disagree with nidx cannot reach into a neighbour's indices (a
    // decreasing pair gives k < 1: BADK). Malformed requests are flagged as
    // a whole by the scatter.
    const int64_t o = (r.off && l > 0) ? r.off[l] : l * (int64_t)r.kuni;
    off[s] = r.idx_base + min(max(o, (int64_t)0), r.nidx);
  }
}

// request-local offsets that do not run from 0 to nidx
__device__ __forceinline__ bool malformed(const ReqEntry &r) {
  return r.off && (r.off[0] != 0 || r.off[r.nsub] != r.nidx);
}

__global__ void scatter_kernel(const ReqEntry *t, int nreq, int64_t tot_sub, const float *lmin,
                               const float *lmax, const float *delta, const uint32_t *status) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < tot_sub; s += stride) {
    const ReqEntry &r = t[find_entry(t, nreq, s, true)];
    const int64_t l = s - r.sub_base;
    if (malformed(r)) {
      r.lmin[l] = r.lmax[l] = r.delta[l] = NAN;
      r.status[l] = pl::ST_BADK;
      continue;
    }
    r.lmin[l] = lmin[s];
    r.lmax[l] = lmax[s];
    r.delta[l] = delta[s];
    r.status[l] = status[s];
  }
}

struct Request {
  int64_t ticket = 0;
  ReqEntry e{};
  hipEvent_t ready = nullptr;  // recorded on the submitter's stream: inputs written
  double work = 0.0;
  Clock::time_point t_submit, t_start, t_done;
  bool done = false;
  pl_status result = PL_OK;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  int32_t *idx = nullptr;
  int64_t *off = nullptr;
  float *out = nullptr;  // lmin | lmax | delta | status, max_subsets each
  void *ws = nullptr;
  ReqEntry *table_dev = nullptr;
  ReqEntry *table_host = nullptr;  // pinned
  std::vector<std::shared_ptr<Request>> reqs;
  int64_t tot_sub = 0;
  bool busy = false;
  pl_status launch_err = PL_OK;
};

double work_of(int form, int kmax, int64_t nsub) {
  return (double)nsub * (double)pl::lds_bytes_of(form, kmax) * (double)kmax;
}

}  // namespace

struct pl_server {
  pl_server_config cfg{};
  const float *G = nullptr;
  int64_t n = 0, ld = 0;
  int device = 0;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_wait;
  std::deque<std::shared_ptr<Request>> queue;
  std::deque<int> inflight;
  std::map<int64_t, std::shared_ptr<Request>> live;
  std::thread sched, comp;
  bool stop = false, sched_exited = false;
  int64_t next_ticket = 1;
  // statistics since the last reset
  int64_t st_requests = 0, st_subsets = 0, st_steps = 0;
  std::vector<double> latencies;
  Clock::time_point first_submit, last_done;
  bool have_first = false;

  void scheduler();
  void completer();
  pl_status launch(Slot &sl);
};

pl_status pl_server::launch(Slot &sl) {
  const int nreq = (int)sl.reqs.size();
  int64_t tot_idx = 0;
  for (int r = 0; r < nreq; ++r) {
    sl.table_host[r] = sl.reqs[r]->e;
    tot_idx = std::max(tot_idx, sl.reqs[r]->e.idx_base + sl.reqs[r]->e.nidx);
  }
  hipError_t e = hipSuccess;
  // the step reads each request's inputs only after the work that wrote them
  for (int r = 0; r < nreq && e == hipSuccess; ++r)
    e = hipStreamWaitEvent(sl.stream, sl.reqs[r]->ready, 0);
  if (e != hipSuccess) return pl::set_hip_error(e, "wait for request inputs");
  e = hipMemcpyAsync(sl.table_dev, sl.table_host, sizeof(ReqEntry) * nreq, hipMemcpyHostToDevice,
                     sl.stream);
  if (e != hipSuccess) return pl::set_hip_error(e, "request table upload");
  const int64_t work = std::max(tot_idx, sl.tot_sub + 1);
  const unsigned blocks = (unsigned)std::min<int64_t>((work + 255) / 256, 4096);
  hipLaunchKernelGGL(assemble_kernel, dim3(blocks), dim3(256), 0, sl.stream, sl.table_dev, nreq,
                     sl.tot_sub, tot_idx, sl.idx, sl.off);
  if ((e = hipGetLastError()) != hipSuccess) return pl::set_hip_error(e, "assemble launch");
  const int64_t cap = cfg.max_subsets;
  float *lmin = sl.out, *lmax = sl.out + cap, *delta = sl.out + 2 * cap;
  uint32_t *status = (uint32_t *)(sl.out + 3 * cap);
  pl::SolveJob j{G,   n,     ld,     sl.idx, sl.off, cfg.max_k, sl.tot_sub,
                 lmin, lmax, delta, status, cfg.form, {0, 0, false}};
  // the steps in flight already overlap on their own streams; helper streams
  // per step would share the hardware queues with them (measured slower)
  j.overlap_bins = false;
  pl_status st = pl::solve_job(j, sl.ws, sl.stream);
  if (st != PL_OK) return st;
  const unsigned sblocks = (unsigned)std::min<int64_t>((sl.tot_sub + 255) / 256, 4096);
  hipLaunchKernelGGL(scatter_kernel, dim3(sblocks), dim3(256), 0, sl.stream, sl.table_dev, nreq,
                     sl.tot_sub, lmin, lmax, delta, status);
  if ((e = hipGetLastError()) != hipSuccess) return pl::set_hip_error(e, "scatter launch");
  if ((e = hipEventRecord(sl.ev, sl.stream)) != hipSuccess)
    return pl::set_hip_error(e, "step event");
  return PL_OK;
}

void pl_server::scheduler() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    int free_slot = -1;
    cv_sched.wait(lk, [&] {
      free_slot = -1;
      for (size_t i = 0; i < slots.size(); ++i)
        if (!slots[i].busy) {
          free_slot = (int)i;
          break;
        }
      return (stop && queue.empty()) || (!queue.empty() && free_slot >= 0);
    });
    if (queue.empty()) break;  // stopping, nothing left to schedule
    Slot &sl = slots[free_slot];
    // pack the head of the queue (arrival order, no overtaking)
    int64_t tot_sub = 0, tot_idx = 0;
    double work = 0.0;
    sl.reqs.clear();
    while (!queue.empty()) {
      const auto &r = queue.front();
      const bool first = sl.reqs.empty();
      if (!first && (tot_sub + r->e.nsub > cfg.max_subsets ||
                     tot_idx + r->e.nidx > cfg.max_indices ||
                     (int)sl.reqs.size() >= cfg.max_requests ||
                     (cfg.work_budget > 0 && work + r->work > cfg.work_budget)))
        break;
      r->e.sub_base = tot_sub;
      r->e.idx_base = tot_idx;
      tot_sub += r->e.nsub;
      tot_idx += r->e.nidx;
      work += r->work;
      sl.reqs.push_back(r);
      queue.pop_front();
    }
    sl.tot_sub = tot_sub;
    sl.busy = true;
    const auto now = Clock::now();
    for (auto &r : sl.reqs) r->t_start = now;
    lk.unlock();
    sl.launch_err = launch(sl);
    lk.lock();
    inflight.push_back(free_slot);
    cv_done.notify_all();
  }
  sched_exited = true;
  cv_done.notify_all();
}

void pl_server::completer() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_done.wait(lk, [&] { return !inflight.empty() || sched_exited; });
    if (inflight.empty()) break;  // the scheduler has exited and nothing is in flight
    // retire whichever in-flight step completes first
    int done_slot = -1;
    hipError_t err = hipSuccess;
    while (done_slot < 0) {
      const std::deque<int> cur = inflight;
      lk.unlock();
      for (int i : cur) {
        Slot &sl = slots[i];
        if (sl.launch_err != PL_OK) {
          (void)hipStreamSynchronize(sl.stream);
          done_slot = i;
          break;
        }
        const hipError_t q = hipEventQuery(sl.ev);
        if (q == hipErrorNotReady) continue;
        err = q;
        done_slot = i;
        break;
      }
      if (done_slot < 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    }
    const auto now = Clock::now();
    Slot &sl = slots[done_slot];
    inflight.erase(std::find(inflight.begin(), inflight.end(), done_slot));
    const pl_status st = sl.launch_err != PL_OK ? sl.launch_err
                         : err != hipSuccess     ? PL_ERR_HIP
                                                 : PL_OK;
    for (auto &r : sl.reqs) {
      if (r->ready) (void)hipEventDestroy(r->ready);
      r->ready = nullptr;
      r->t_done = now;
      r->done = true;
      r->result = st;
      latencies.push_back(std::chrono::duration<double>(now - r->t_submit).count());
      st_subsets += r->e.nsub;
    }
    st_requests += (int64_t)sl.reqs.size();
    st_steps += 1;
    last_done = now;
    sl.reqs.clear();
    sl.busy = false;
    cv_wait.notify_all();
    cv_sched.notify_all();
  }
}

extern "C" {

pl_status pl_server_create(const float *G, int64_t n, int64_t ld, const pl_server_config *cfg,
                           pl_server **out) {
  if (!out || !cfg || !G) return pl::set_error(PL_ERR_ARG, "null argument");
  *out = nullptr;
  if (n < 1 || ld < n) return pl::set_error(PL_ERR_ARG, "bad n or ld");
  if (cfg->max_subsets < 1 || cfg->max_subsets > 0x3fffffffLL || cfg->max_indices < 1 ||
      cfg->max_requests < 1 || cfg->max_k < 1 || cfg->max_k > pl_max_k() || cfg->depth < 1 ||
      cfg->depth > 8)
    return pl::set_error(PL_ERR_ARG, "server configuration out of range");
  pl::SolveJob probe{G, n, ld, (const int32_t *)G, (const int64_t *)G, cfg->max_k, 1,
                     (float *)G, (float *)G, (float *)G, (uint32_t *)G, cfg->form, {0, 0, false}};
  pl_status st = pl::validate_job(probe);
  if (st != PL_OK) return st;
  auto *s = new pl_server;
  s->cfg = *cfg;
  s->G = G;
  s->n = n;
  s->ld = ld;
  hipError_t e = hipGetDevice(&s->device);
  s->slots.resize(cfg->depth);
  const int64_t cap = cfg->max_subsets;
  for (auto &sl : s->slots) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&sl.idx, sizeof(int32_t) * cfg->max_indices);
    if (e == hipSuccess) e = hipMalloc(&sl.off, sizeof(int64_t) * (cap + 1));
    if (e == hipSuccess) e = hipMalloc(&sl.out, sizeof(float) * 4 * cap);
    if (e == hipSuccess) e = hipMalloc(&sl.ws, pl::workspace_bytes(cap, cfg->max_k));
    if (e == hipSuccess) e = hipMalloc(&sl.table_dev, sizeof(ReqEntry) * cfg->max_requests);
    if (e == hipSuccess) e = hipHostMalloc(&sl.table_host, sizeof(ReqEntry) * cfg->max_requests);
  }
  if (e != hipSuccess) {
    st = pl::set_hip_error(e, "server allocation");
    pl_server_destroy(s);
    return st;
  }
  s->sched = std::thread([s] { s->scheduler(); });
  s->comp = std::thread([s] { s->completer(); });
  *out = s;
  return PL_OK;
}

pl_status pl_server_submit(pl_server *s, const int32_t *idx, const int64_t *offsets, int k,
                           int64_t nsub, int64_t nidx, int kmax, float *lmin, float *lmax,
                           float *delta, uint32_t *status, void *stream, int64_t *ticket) {
  if (!s || !idx || !lmin || !lmax || !delta || !status || !ticket)
    return pl::set_error(PL_ERR_ARG, "null argument");
  if (nsub < 1 || nsub > s->cfg.max_subsets)
    return pl::set_error(PL_ERR_ARG, "nsub outside [1, max_subsets]");
  if (nidx < 1 || nidx > s->cfg.max_indices)
    return pl::set_error(PL_ERR_ARG, "nidx outside [1, max_indices]");
  if (!offsets) {
    if (k < 1 || k > s->cfg.max_k || nidx != nsub * (int64_t)k)
      return pl::set_error(PL_ERR_ARG, "uniform request needs 1 <= k <= max_k, nidx = nsub*k");
    kmax = k;
  } else if (kmax < 1 || kmax > s->cfg.max_k) {
    return pl::set_error(PL_ERR_ARG, "kmax outside [1, max_k]");
  }
  auto r = std::make_shared<Request>();
  r->e.idx = idx;
  r->e.off = offsets;
  r->e.nsub = nsub;
  r->e.nidx = nidx;
  r->e.kuni = offsets ? 0 : k;
  r->e.lmin = lmin;
  r->e.lmax = lmax;
  r->e.delta = delta;
  r->e.status = status;
  r->work = work_of(s->cfg.form, kmax, nsub);
  hipError_t e = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(r->ready, (hipStream_t)stream);
  if (e != hipSuccess) {
    if (r->ready) (void)hipEventDestroy(r->ready);
    r->ready = nullptr;
    return pl::set_hip_error(e, "request event");
  }
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      (void)hipEventDestroy(r->ready);
      r->ready = nullptr;
      return pl::set_error(PL_ERR_ARG, "server is shutting down");
    }
    r->ticket = s->next_ticket++;
    if (!s->have_first) {
      s->first_submit = r->t_submit;
      s->have_first = true;
    }
    s->live[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
  }
  s->cv_sched.notify_all();
  return PL_OK;
}

pl_status pl_server_wait(pl_server *s, int64_t ticket, double *latency_s, double *queued_s) {
  if (!s) return pl::set_error(PL_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return pl::set_error(PL_ERR_ARG, "unknown or already retired ticket");
  auto r = it->second;
  s->cv_wait.wait(lk, [&] { return r->done; });
  s->live.erase(ticket);
  if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
  if (queued_s) *queued_s = std::chrono::duration<double>(r->t_start - r->t_submit).count();
  if (r->result != PL_OK) return pl::set_error(r->result, "request failed on the device");
  return PL_OK;
}

pl_status pl_server_poll(pl_server *s, int64_t ticket, int *done) {
  if (!s || !done) return pl::set_error(PL_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return pl::set_error(PL_ERR_ARG, "unknown or already retired ticket");
  *done = it->second->done ? 1 : 0;
  return PL_OK;
}

pl_status pl_server_query(pl_server *s, pl_server_stats *out, int reset) {
  if (!s || !out) return pl::set_error(PL_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  out->requests = s->st_requests;
  out->subsets = s->st_subsets;
  out->steps = s->st_steps;
  out->span_s = s->have_first && s->st_requests
                    ? std::chrono::duration<double>(s->last_done - s->first_submit).count()
                    : 0.0;
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    const size_t i = (size_t)std::min<double>(l.size() - 1, std::ceil(p * l.size()) - 1);
    return l[i];
  };
  out->p50_s = pct(0.50);
  out->p99_s = pct(0.99);
  out->max_s = l.empty() ? 0.0 : l.back();
  if (reset) {
    s->st_requests = s->st_subsets = s->st_steps = 0;
    s->latencies.clear();
    s->have_first = false;
  }
  return PL_OK;
}

pl_status pl_server_destroy(pl_server *s) {
  if (!s) return PL_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_sched.notify_all();
  s->cv_done.notify_all();
  if (s->sched.joinable()) s->sched.join();
  else s->sched_exited = true;
  s->cv_done.notify_all();
  if (s->comp.joinable()) s->comp.join();
  for (auto &sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.table_host) (void)hipHostFree(sl.table_host);
    if (sl.table_dev) (void)hipFree(sl.table_dev);
    if (sl.ws) (void)hipFree(sl.ws);
    if (sl.out) (void)hipFree(sl.out);
    if (sl.off) (void)hipFree(sl.off);
    if (sl.idx) (void)hipFree(sl.idx);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  delete s;
  return PL_OK;
}

}  // extern "C"
