// C ABI of libparalanczos.so: the device-pointer API of paralanczos.h and the
// drop-in host entry points of ParaLanczos.h. Host code only; kernels live in
// pl_kernels.hip.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <functional>
#include <iterator>
#include <map>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "../../include/ParaLanczos.h"
#include "../../include/paralanczos.h"
#include "pl_common.hpp"
#include "pl_internal.hpp"

namespace {

thread_local std::string g_err;

pl_status fail(pl_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

pl_status hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? PL_ERR_NOMEM : PL_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
}

constexpr int kMaxK = pl::KMAX_ALL;

// lanes per subset of the default variant for k: the smallest lane group of
// 8..64 lanes (one row per lane) that takes k, beyond k = 64 the half-row
// kernels (256 lanes up to k = 128, 512 lanes up to 256)
int group_for(int k) {
  if (k > 128) return 512;
  if (k > 64) return 256;
  int g = 8;
  while (g < k) g <<= 1;
  return g;
}

// ---------------------------------------------------------------------------
// Tuning table: the solve variant (lanes per subset, LDS row padding) used for
// each k by the tuned form. Defaults: group_for(k) lanes, pad 4; pl_autotune
// overwrites entries with measured winners.
// ---------------------------------------------------------------------------
using pl::Variant;
std::mutex g_tune_mu;
Variant g_tune[pl::KMAX_ALL + 1];
bool g_tune_ready = false;

Variant default_variant(int k) { return Variant{group_for(k), 4, false}; }

// kernel family of a variant: results are bit-identical within a family
int family(int lanes) { return lanes < 128 ? 0 : lanes; }

constexpr int kFormMask = 0xff;
constexpr int kPrecMask = PL_FORM_REORTH_F32 | PL_FORM_REORTH_F64;

// FP64 Gram-Schmidt coefficients: forced by a flag, otherwise the tuned form
// uses them from k = 128 on (variants of 128 lanes and more carry them)
bool use_f64(int form, int k, int lanes) {
  if (lanes < 128) return false;
  if (form & PL_FORM_REORTH_F64) return true;
  if (form & PL_FORM_REORTH_F32) return false;
  return (form & kFormMask) == PL_FORM_TUNED && k >= 128;
}

void tune_init_locked() {
  if (g_tune_ready) return;
  for (int k = 1; k <= pl::KMAX_ALL; ++k) g_tune[k] = default_variant(k);
  g_tune_ready = true;
}

Variant tuned_variant(int k) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_init_locked();
  return g_tune[k];
}

Variant plain_variant(int k) {
  // one 64-lane wavefront per subset; a 128-lane workgroup beyond k = 64 and
  // a 512-lane one beyond k = 128
  return Variant{k <= 64 ? 64 : k <= 128 ? 128 : 512, 4, false};
}

Variant plain_wave_variant(int k) {
  // one 64-lane wavefront per subset at every k
  return k <= 64 ? Variant{64, 4, false} : Variant{64, 0, false, true};
}

bool is_plain(int form) {
  return (form & kFormMask) == PL_FORM_PLAIN || (form & kFormMask) == PL_FORM_PLAIN_WAVE;
}

// Bins of a mixed-size batch. Tuned: maximal runs of consecutive k sharing a
// variant (the bin edges follow the tuning table). Plain: one bin per exact k.
// Bin 0 also takes every size outside [1, kmax]; its launch flags them BADK.
struct BinPlan {
  pl::BinLut lut;
  int nbins = 0;
  Variant var[pl::MAXBIN];
  int kbin[pl::MAXBIN];  // plain: the k of the bin (0 for the invalid sizes)
};

Variant job_variant(int form, int k) {
  Variant v = (form & kFormMask) == PL_FORM_PLAIN        ? plain_variant(k)
              : (form & kFormMask) == PL_FORM_PLAIN_WAVE ? plain_wave_variant(k)
                                                         : tuned_variant(k);
  v.f64 = use_f64(form, k, v.lanes);
  return v;
}

void make_bins(int form, int kmax, BinPlan &bp) {
  bp.nbins = 0;
  if (is_plain(form)) {
    bp.var[0] = job_variant(form, 1);
    bp.kbin[0] = 0;
    for (int k = 1; k <= kmax; ++k) {
      bp.var[k] = job_variant(form, k);
      bp.kbin[k] = k;
    }
    bp.nbins = kmax + 1;
    for (int i = 0; i < pl::KLUT; ++i) bp.lut.bin[i] = (i >= 1 && i <= kmax) ? (uint16_t)i : 0;
    return;
  }
  int prev_l = -1, prev_p = -1, prev_f = -1;
  uint16_t cur = 0;
  for (int i = 0; i < pl::KLUT; ++i) bp.lut.bin[i] = 0;
  for (int k = 1; k <= kmax; ++k) {
    const Variant v = job_variant(form, k);
    if (v.lanes != prev_l || v.pad != prev_p || (int)v.f64 != prev_f) {
      cur = (uint16_t)bp.nbins++;
      bp.var[cur] = v;
      bp.kbin[cur] = k;
      prev_l = v.lanes;
      prev_p = v.pad;
      prev_f = v.f64;
    }
    bp.lut.bin[k] = cur;
  }
}

pl_status check_solve_args(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                           int64_t nsub) {
  // an empty batch reads and writes nothing: idx may then be null
  if (!G || (!idx && nsub > 0)) return fail(PL_ERR_ARG, "G and idx must be non-null");
  if (n < 1 || n > 0x7fffffffLL) return fail(PL_ERR_ARG, "n = %lld out of range", (long long)n);
  if (ld < n) return fail(PL_ERR_ARG, "ld = %lld < n = %lld", (long long)ld, (long long)n);
  if (nsub < 0 || nsub > 0x7fffffffLL)
    return fail(PL_ERR_ARG, "nsub = %lld outside [0, 2^31)", (long long)nsub);
  if (k < 1) return fail(PL_ERR_ARG, "k = %d < 1", k);
  if (k > kMaxK) return fail(PL_ERR_UNSUPPORTED, "k = %d > %d is not supported", k, kMaxK);
  return PL_OK;
}

// stream-ordered workspace: the caller's, or one allocated for this call
struct Workspace {
  void *p = nullptr;
  bool owned = false;
  hipStream_t stream = nullptr;
  pl_status acquire(void *given, size_t given_bytes, size_t need, hipStream_t s) {
    stream = s;
    if (given) {
      if (given_bytes < need)
        return fail(PL_ERR_ARG, "workspace of %zu bytes < %zu needed", given_bytes, need);
      p = given;
      return PL_OK;
    }
    hipError_t e = hipMallocAsync(&p, need ? need : 16, s);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(workspace)");
    owned = true;
    return PL_OK;
  }
  ~Workspace() {
    if (owned) (void)hipFreeAsync(p, stream);
  }
};

using pl::SolveJob;

pl::SolveArgs solve_args(const SolveJob &j, const pl::WorkspaceParts &w) {
  pl::SolveArgs a{};
  a.G = j.G;
  a.n = j.n;
  a.ld = j.ld;
  a.idx = j.idx;
  a.offsets = j.offsets;
  a.kuni = j.offsets ? 0 : j.k;
  a.kcap = j.k;
  a.nslots = j.nsub;
  a.status = j.status;
  a.tri = w.tri;
  a.tstride = j.k;
  a.bounds = w.bounds;
  return a;
}

pl_status run_bisect(const SolveJob &j, const pl::WorkspaceParts &w, const int32_t *order,
                     hipStream_t s) {
  pl::BisectArgs b{};
  b.order = order;
  b.tri = w.tri;
  b.tstride = j.k;
  b.bounds = w.bounds;
  b.offsets = j.offsets;
  b.kuni = j.offsets ? 0 : j.k;
  b.nsub = j.nsub;
  b.status = j.status;
  b.lmin = j.lmin;
  b.lmax = j.lmax;
  b.delta = j.delta;
  hipError_t e = pl::launch_bisect(0, b, s);
  if (e != hipSuccess) return hip_fail(e, "bisection launch");
  return PL_OK;
}

// Whether a job runs without host synchronisation (and so can be captured
// into a graph): all but the plain form of a mixed-size batch.
bool job_is_async(const SolveJob &j) {
  if (j.form != -1 && (j.form & kFormMask) == PL_FORM_PLAIN_WAVE) return false;
  return !(j.offsets && j.form != -1 && (j.form & kFormMask) == PL_FORM_PLAIN);
}

// one launch of a variant over the job's slots; the one-wavefront variant
// on the job's scratch, or on one allocated on the stream for the launch
hipError_t launch_variant(const Variant &v, const pl::SolveArgs &a, hipStream_t s,
                          int64_t max_slots, const SolveJob &j) {
  if (!v.wave) return pl::launch_solve(v.lanes, v.pad, pl::MODE_TRIDIAG, a, s, max_slots, v.f64);
  const int64_t want = a.bins ? max_slots : a.nslots;
  if (j.wave_scratch)
    return pl::launch_wave(a, j.wave_scratch,
                           std::max<int64_t>(1, std::min<int64_t>(want, j.wave_slots)), s);
  const int64_t slots = std::max<int64_t>(1, std::min<int64_t>(want, pl::kWaveSlots));
  void *scratch = nullptr;
  hipError_t e = hipMallocAsync(&scratch, pl::wave_scratch_bytes(slots), s);
  if (e != hipSuccess) return e;
  e = pl::launch_wave(a, (float *)scratch, slots, s);
  const hipError_t f = hipFreeAsync(scratch, s);
  return e != hipSuccess ? e : f;
}

// The stream-ordered work of one solve on a workspace of
// pl::workspace_bytes(nsub, k) bytes.
//   uniform k: one launch of the form's variant, then the bisection;
//   mixed, tuned: device-side binning by variant, one persistent launch per
//     bin draining its run of the order array, bisection in bin order;
//   mixed, plain: binning by exact k, bin counts to the host, one launch per
//     k sized to its count, bisection in bin order.
constexpr int kStageTridiag = 1, kStageBisect = 2;

// Helper streams of mixed tuned batches: a fixed pool of kHelpers streams
// per device (default priority, created on first use with the calling
// stream's device current, kept until the process ends); a calling stream
// takes the pool entry its handle hashes to, so the number of helper streams
// stays bounded however many calling streams come and go. A mixed tuned
// batch runs its widest bin (512 lanes: one workgroup fills a CU) on the
// caller's stream and the other bins on the helper: their workgroups cannot
// fit beside the wide kernel's, so they take CUs as its persistent
// workgroups drain and fill its tail instead of waiting behind it. Batches
// of two calling streams that share a helper have their small bins ordered
// after one another on it.
constexpr int kHelpers = 2;
std::mutex g_helper_mu;
std::map<int, std::vector<hipStream_t>> g_helper;
std::atomic<int64_t> g_forks{0};

hipError_t helper_for(hipStream_t s, hipStream_t *h) {
  hipDevice_t sdev = 0;
  hipError_t e = hipStreamGetDevice(s, &sdev);
  if (e != hipSuccess) return e;
  const int dev = (int)sdev;
  const size_t slot = (std::hash<const void *>{}((const void *)s) >> 4) % kHelpers;
  std::lock_guard<std::mutex> lock(g_helper_mu);
  std::vector<hipStream_t> &pool = g_helper[dev];
  if (pool.empty()) pool.assign(kHelpers, nullptr);
  if (pool[slot]) {
    *h = pool[slot];
    return hipSuccess;
  }
  int cur = 0;
  e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
  e = hipStreamCreateWithFlags(h, hipStreamNonBlocking);
  if (cur != dev) (void)hipSetDevice(cur);
  if (e == hipSuccess) pool[slot] = *h;
  return e;
}

// the tuned bins of a mixed batch: the widest on s, the rest on the helper
// stream between a fork after the binning and a join before the bisection
// (not under graph capture, where everything stays on s). Once the helper
// waits on the fork, s always waits on the helper's join, also when a launch
// failed, so nothing queued on the helper outlives what s orders after it.
hipError_t launch_tuned_bins(const BinPlan &bp, pl::SolveArgs a, hipStream_t s, int64_t nsub,
                             bool overlap) {
  int wide = -1;
  for (int b = 0; b < bp.nbins; ++b)
    if (bp.var[b].lanes == 512) wide = b;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(s, &cs);
  if (e != hipSuccess) return e;
  hipStream_t h = nullptr;
  hipEvent_t binned = nullptr, rest = nullptr;
  bool fork = overlap && wide >= 0 && bp.nbins > 1 && cs == hipStreamCaptureStatusNone &&
              helper_for(s, &h) == hipSuccess;
  if (fork) {
    fork = hipEventCreateWithFlags(&binned, hipEventDisableTiming) == hipSuccess &&
           hipEventCreateWithFlags(&rest, hipEventDisableTiming) == hipSuccess &&
           hipEventRecord(binned, s) == hipSuccess && hipStreamWaitEvent(h, binned, 0) == hipSuccess;
    if (!fork) {
      // the helper may wait on `binned` already: it orders nothing after
      // s's binning, and no work goes to the helper
      if (binned) (void)hipEventDestroy(binned);
      if (rest) (void)hipEventDestroy(rest);
    }
  }
  if (!fork) {
    for (int b = 0; b < bp.nbins; ++b) {
      a.bin = b;
      e = pl::launch_solve(bp.var[b].lanes, bp.var[b].pad, pl::MODE_TRIDIAG, a, s, nsub,
                           bp.var[b].f64);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  g_forks.fetch_add(1, std::memory_order_relaxed);
  a.bin = wide;
  e = pl::launch_solve(bp.var[wide].lanes, bp.var[wide].pad, pl::MODE_TRIDIAG, a, s, nsub,
                       bp.var[wide].f64);
  for (int b = 0; b < bp.nbins && e == hipSuccess; ++b) {
    if (b == wide) continue;
    a.bin = b;
    e = pl::launch_solve(bp.var[b].lanes, bp.var[b].pad, pl::MODE_TRIDIAG, a, h, nsub,
                         bp.var[b].f64);
  }
  // the join, unconditionally
  hipError_t f = hipEventRecord(rest, h);
  if (f == hipSuccess) f = hipStreamWaitEvent(s, rest, 0);
  if (f != hipSuccess) f = hipStreamSynchronize(h);  // no event: wait for the helper
  // destroying a recorded event releases it once it completes
  (void)hipEventDestroy(binned);
  (void)hipEventDestroy(rest);
  return e != hipSuccess ? e : f;
}

pl_status run_job(const SolveJob &j, void *ws, hipStream_t s, int stages = 3) {
  const pl::WorkspaceParts w = pl::split_workspace(ws, j.nsub, j.k);
  pl::SolveArgs a = solve_args(j, w);
  hipError_t e;
  if (!j.offsets || j.form == -1) {
    // uniform size, or an explicit variant over a mixed batch: one launch
    if (stages & kStageTridiag) {
      const Variant v = j.form == -1 ? j.explicit_v : job_variant(j.form, j.k);
      e = launch_variant(v, a, s, 0, j);
      if (e != hipSuccess) return hip_fail(e, "solve launch");
    }
    return (stages & kStageBisect) ? run_bisect(j, w, nullptr, s) : PL_OK;
  }
  if (!(stages & kStageTridiag)) return (stages & kStageBisect) ? run_bisect(j, w, w.order, s) : PL_OK;
  BinPlan bp;
  make_bins(j.form, j.k, bp);
  e = pl::launch_binning(j.offsets, 0, j.nsub, bp.lut, bp.nbins, w.table, w.order, s);
  if (e != hipSuccess) return hip_fail(e, "binning launch");
  a.order = w.order;
  if ((j.form & kFormMask) == PL_FORM_TUNED) {
    a.bins = w.table;
    e = launch_tuned_bins(bp, a, s, j.nsub, j.overlap_bins && !(j.form & PL_FORM_NO_OVERLAP));
    if (e != hipSuccess) return hip_fail(e, "persistent solve launch");
  } else {
    std::array<int32_t, 2 * pl::MAXBIN> tab;
    e = hipMemcpyAsync(tab.data(), w.table, tab.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                       s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "bin counts to host");
    for (int b = 0; b < bp.nbins; ++b) {
      const int cnt = tab[pl::BIN_COUNT + b];
      if (cnt == 0) continue;
      a.order = w.order + tab[pl::BIN_START + b];
      a.nslots = cnt;
      e = launch_variant(bp.var[b], a, s, 0, j);
      if (e != hipSuccess) return hip_fail(e, "per-k solve launch");
    }
  }
  return (stages & kStageBisect) ? run_bisect(j, w, w.order, s) : PL_OK;
}

pl_status check_job(const SolveJob &j) {
  pl_status st = check_solve_args(j.G, j.n, j.ld, j.idx, j.k, j.nsub);
  if (st != PL_OK) return st;
  if (j.nsub > 0 && (!j.lmin || !j.lmax || !j.delta || !j.status))
    return fail(PL_ERR_ARG, "output pointers must be non-null");
  if (j.form == -1) {
    if (pl::variant_capacity(j.explicit_v.lanes) < j.k)
      return fail(PL_ERR_ARG, "lanes per subset %d take k <= %d < k = %d", j.explicit_v.lanes,
                  pl::variant_capacity(j.explicit_v.lanes), j.k);
    if (!pl::solve_variant_exists(j.explicit_v.lanes, j.explicit_v.pad))
      return fail(PL_ERR_ARG, "no variant with %d lanes and pad %d", j.explicit_v.lanes,
                  j.explicit_v.pad);
  } else if (((j.form & kFormMask) != PL_FORM_TUNED && !is_plain(j.form)) ||
             (j.form & ~(kFormMask | kPrecMask | PL_FORM_NO_OVERLAP)) ||
             (j.form & kPrecMask) == kPrecMask) {
    return fail(PL_ERR_ARG, "unknown form %d", j.form);
  } else if ((j.form & kFormMask) == PL_FORM_PLAIN_WAVE && (j.form & kPrecMask)) {
    return fail(PL_ERR_ARG, "the one-wavefront form is FP32 only (no precision flags)");
  }
  return PL_OK;
}

pl_status solve_once(const SolveJob &j, void *ws, size_t ws_bytes, void *stream) {
  pl_status st = check_job(j);
  if (st != PL_OK) return st;
  if (j.nsub == 0) return PL_OK;
  hipStream_t s = (hipStream_t)stream;
  Workspace w;
  st = w.acquire(ws, ws_bytes, pl::workspace_bytes(j.nsub, j.k), s);
  if (st != PL_OK) return st;
  return run_job(j, w.p, s);
}

// ---------------------------------------------------------------------------
// The workspace of pl_solve_noalloc: a file-scope device array (each device
// holds its own copy from module load until the process ends). Host code
// never uses the array's name as an address: the address of each device's
// copy comes from hipGetSymbolAddress, resolved once per device with that
// device current (by pl_init, or by the first pl_solve_noalloc on it).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = (size_t)32 << 20;  // PL_NOALLOC_SCRATCH_BYTES
static_assert(kNoAllocBytes == PL_NOALLOC_SCRATCH_BYTES, "header and library agree");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// this device's copy of `scratch` (the current device)
pl_status noalloc_scratch(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(PL_ERR_UNSUPPORTED, "device %d >= %d devices", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(PL_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PL_OK;
}

}  // namespace

namespace pl {
pl_status solve_job(const SolveJob &j, void *ws, hipStream_t s) { return run_job(j, ws, s); }
pl_status validate_job(const SolveJob &j) { return check_job(j); }
pl_status set_error(pl_status st, const char *msg) { return fail(st, "%s", msg); }
pl_status set_hip_error(hipError_t e, const char *what) { return hip_fail(e, what); }
int lds_bytes_of(int form, int k) {
  const Variant v = job_variant(form, k);
  if (v.lanes == 512) return 159 * 1024;  // the row-part kernels' LDS, to the KiB
  if (v.lanes == 256) return 51 * 1024;
  if (v.lanes == 128) return 128 * (128 + v.pad) * 4;
  return v.lanes * (v.lanes + v.pad) * 4;
}
}  // namespace pl

extern "C" {

const char *pl_version(void) { return "paralanczos-amd 0.1 (gfx950)"; }

const char *pl_status_string(pl_status st) {
  switch (st) {
    case PL_OK: return "ok";
    case PL_ERR_ARG: return "invalid argument";
    case PL_ERR_HIP: return "HIP runtime error";
    case PL_ERR_NOMEM: return "out of device memory";
    case PL_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char *pl_last_error(void) { return g_err.c_str(); }

pl_status pl_init(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  void *p = nullptr;
  return noalloc_scratch(&p);
}

int pl_max_k(void) { return kMaxK; }

int64_t pl_overlap_forks(void) { return g_forks.load(std::memory_order_relaxed); }

size_t pl_solve_workspace_size(int64_t nsub, int k) {
  if (nsub < 0 || k < 1) return 0;
  return pl::workspace_bytes(nsub, k);
}

pl_status pl_solve(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                   const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                   float *delta, uint32_t *status, int form, void *stream) {
  return pl_solve_ws(G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, form, nullptr, 0,
                     stream);
}

pl_status pl_solve_ws(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                      const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                      float *delta, uint32_t *status, int form, void *workspace,
                      size_t workspace_bytes, void *stream) {
  SolveJob j{G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, form, {0, 0, false}};
  return solve_once(j, workspace, workspace_bytes, stream);
}

pl_status pl_solve_noalloc(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                           const int64_t *offsets, int k, int64_t nsub, float *lmin,
                           float *lmax, float *delta, uint32_t *status, int form,
                           void *stream) {
  SolveJob j{G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, form, {0, 0, false}};
  j.overlap_bins = false;  // no helper stream or events to create
  pl_status st = check_job(j);
  if (st != PL_OK) return st;
  if ((form & kFormMask) == PL_FORM_PLAIN_WAVE)
    return fail(PL_ERR_UNSUPPORTED,
                "pl_solve_noalloc: the one-wavefront plain form allocates its own scratch");
  const size_t need = pl::workspace_bytes(nsub, k);
  if (need > kNoAllocBytes)
    return fail(PL_ERR_ARG,
                "pl_solve_noalloc: %lld subsets of k <= %d need %zu bytes of scratch > %zu "
                "(PL_NOALLOC_SCRATCH_BYTES)",
                (long long)nsub, k, need, kNoAllocBytes);
  void *ws = nullptr;
  st = noalloc_scratch(&ws);
  if (st != PL_OK) return st;
  if (nsub == 0) return PL_OK;
  return run_job(j, ws, (hipStream_t)stream);
}

pl_status pl_upper(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                   const int64_t *offsets, int k, int64_t nsub, int iters, float *upper,
                   uint32_t *status, void *stream) {
  pl_status st = check_solve_args(G, n, ld, idx, k, nsub);
  if (st != PL_OK) return st;
  if (nsub > 0 && (!upper || !status)) return fail(PL_ERR_ARG, "output pointers must be non-null");
  if (iters < 1 || iters > 1000000)
    return fail(PL_ERR_ARG, "iters = %d outside [1, 1000000]", iters);
  if (nsub == 0) return PL_OK;
  pl::SolveArgs a{};
  a.G = G;
  a.n = n;
  a.ld = ld;
  a.idx = idx;
  a.offsets = offsets;
  a.kuni = offsets ? 0 : k;
  a.kcap = k;
  a.nslots = nsub;
  a.status = status;
  hipError_t e = pl::launch_upper(a, iters, upper, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "power-iteration launch");
  return PL_OK;
}

pl_status pl_solve_stage(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                         const int64_t *offsets, int k, int64_t nsub, float *lmin, float *lmax,
                         float *delta, uint32_t *status, int form, int stage, void *workspace,
                         size_t workspace_bytes, void *stream) {
  SolveJob j{G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, form, {0, 0, false}};
  pl_status st = check_job(j);
  if (st != PL_OK) return st;
  if (stage != kStageTridiag && stage != kStageBisect)
    return fail(PL_ERR_ARG, "stage must be 1 (tridiagonalise) or 2 (bisect)");
  if (!workspace) return fail(PL_ERR_ARG, "pl_solve_stage needs a caller workspace");
  if (workspace_bytes < pl::workspace_bytes(nsub, k))
    return fail(PL_ERR_ARG, "workspace of %zu bytes < %zu needed", workspace_bytes,
                pl::workspace_bytes(nsub, k));
  if (nsub == 0) return PL_OK;
  return run_job(j, workspace, (hipStream_t)stream, stage);
}

pl_status pl_solve_variant(const float *G, int64_t n, int64_t ld, const int32_t *idx,
                           const int64_t *offsets, int k, int64_t nsub, float *lmin,
                           float *lmax, float *delta, uint32_t *status, int lanes, int pad,
                           void *stream) {
  SolveJob j{G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, -1, {lanes, pad, false}};
  return solve_once(j, nullptr, 0, stream);
}

// ---------------------------------------------------------------------------
// Plans: a solve bound to fixed device buffers, with its own workspace. The
// first execution of an asynchronous plan records its launch sequence into a
// HIP graph; later executions replay the graph (the bins are recomputed on the
// device from the offsets' current contents each time).
// ---------------------------------------------------------------------------
}  // extern "C"

struct pl_plan {
  SolveJob job;
  void *ws = nullptr;
  void *wave = nullptr;  // the one-wavefront form's scratch
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::mutex mu;
};

extern "C" {

pl_status pl_plan_create(pl_plan **plan, const float *G, int64_t n, int64_t ld,
                         const int32_t *idx, const int64_t *offsets, int k, int64_t nsub,
                         float *lmin, float *lmax, float *delta, uint32_t *status, int form) {
  if (!plan) return fail(PL_ERR_ARG, "plan must be non-null");
  *plan = nullptr;
  SolveJob j{G, n, ld, idx, offsets, k, nsub, lmin, lmax, delta, status, form, {0, 0, false}};
  pl_status st = check_job(j);
  if (st != PL_OK) return st;
  pl_plan *p = new pl_plan;
  p->job = j;
  hipError_t e = hipMalloc(&p->ws, pl::workspace_bytes(nsub > 0 ? nsub : 1, k));
  if (e == hipSuccess && (form & kFormMask) == PL_FORM_PLAIN_WAVE && k > 64) {
    // resident wavefronts of the one-wavefront form: its slots, at most kWaveSlots
    const int64_t slots = std::max<int64_t>(1, std::min<int64_t>(nsub, pl::kWaveSlots));
    e = hipMalloc(&p->wave, pl::wave_scratch_bytes(slots));
    p->job.wave_scratch = (float *)p->wave;
    p->job.wave_slots = slots;
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    if (p->ws) (void)hipFree(p->ws);
    if (p->wave) (void)hipFree(p->wave);
    delete p;
    return hip_fail(e, "plan allocation");
  }
  *plan = p;
  return PL_OK;
}

pl_status pl_plan_execute(pl_plan *p, void *stream) {
  if (!p) return fail(PL_ERR_ARG, "plan must be non-null");
  std::lock_guard<std::mutex> lock(p->mu);
  hipStream_t s = (hipStream_t)stream;
  if (p->job.nsub == 0) return PL_OK;
  if (!job_is_async(p->job)) return run_job(p->job, p->ws, s);
  if (!p->exec) {
    hipError_t e = hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    pl_status st = run_job(p->job, p->ws, p->cap);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(p->cap, &g);
    if (st != PL_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      p->exec = nullptr;
      return hip_fail(e, "hipGraphInstantiate");
    }
    p->graph = g;
  }
  hipError_t e = hipGraphLaunch(p->exec, s);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  return PL_OK;
}

int pl_plan_is_graph(const pl_plan *p) { return p && p->exec ? 1 : 0; }

pl_status pl_plan_destroy(pl_plan *p) {
  if (!p) return PL_OK;
  hipError_t e = hipSuccess;
  if (p->cap) e = hipStreamSynchronize(p->cap);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  // the plan's workspace may still be in use by a replay on the caller's
  // stream: wait for the device before releasing it
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (p->ws) (void)hipFree(p->ws);
  if (p->wave) (void)hipFree(p->wave);
  delete p;
  if (e != hipSuccess) return hip_fail(e, "plan teardown");
  return PL_OK;
}

// ---------------------------------------------------------------------------
// Tuning table access and the autotuner
// ---------------------------------------------------------------------------
pl_status pl_tuning_get(int k, int *lanes, int *pad) {
  if (k < 1 || k > kMaxK || !lanes || !pad) return fail(PL_ERR_ARG, "bad k or null output");
  const Variant v = tuned_variant(k);
  *lanes = v.lanes;
  *pad = v.pad;
  return PL_OK;
}

pl_status pl_tuning_set(int k, int lanes, int pad) {
  if (k < 1 || k > kMaxK) return fail(PL_ERR_ARG, "k = %d outside [1, %d]", k, kMaxK);
  if (!pl::solve_variant_exists(lanes, pad) || pl::variant_capacity(lanes) < k)
    return fail(PL_ERR_ARG, "no variant with %d lanes and pad %d for k = %d", lanes, pad, k);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_init_locked();
  g_tune[k] = Variant{lanes, pad, false};
  return PL_OK;
}

pl_status pl_tuning_reset(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune_ready = false;
  tune_init_locked();
  return PL_OK;
}

}  // extern "C"

namespace {
// The autotuner. Per listed k it times the candidate variants on `sample`
// seeded subsets of G (three runs, the first a warm-up, the best of the other
// two) and keeps the fastest. Candidates: the variants of the family of k's
// default, or with `edges` every variant that takes k, so that the family
// boundaries (the bin edges of mixed batches) move to the measured optimum.
// sweep (nk x PL_TUNE_VARIANTS, may be null) gets every candidate's time.
pl_status autotune(const float *G, int64_t n, int64_t ld, const int32_t *klist, int nk,
                   int64_t sample, uint64_t seed, bool edges, double *best_ms, double *sweep,
                   void *stream) {
  if (!G || !klist || nk < 1) return fail(PL_ERR_ARG, "G and klist must be non-null, nk >= 1");
  if (sample < 1 || sample > (1 << 24)) return fail(PL_ERR_ARG, "sample outside [1, 2^24]");
  int kmax = 0;
  for (int i = 0; i < nk; ++i) {
    if (klist[i] < 1 || klist[i] > kMaxK || klist[i] > n)
      return fail(PL_ERR_ARG, "klist[%d] = %d outside [1, min(n, %d)]", i, klist[i], kMaxK);
    kmax = std::max(kmax, klist[i]);
  }
  hipStream_t s = (hipStream_t)stream;
  int32_t *idx = nullptr;
  float *out = nullptr;
  void *ws = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const size_t wsb = pl::workspace_bytes(sample, kmax);
  hipError_t e = hipMalloc(&idx, (size_t)sample * kmax * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&out, (size_t)sample * 4 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&ws, wsb);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  pl_status st = e == hipSuccess ? PL_OK : hip_fail(e, "autotune allocation");
  static const int kLanes[] = {8, 16, 32, 64, 128, 256, 512};
  static const int kPads[] = {0, 1, 4};
  static_assert(sizeof(kLanes) / sizeof(int) * 3 == PL_TUNE_VARIANTS, "variant grid");
  if (sweep)
    for (int v = 0; v < nk * PL_TUNE_VARIANTS; ++v) sweep[v] = -1.0;
  for (int i = 0; i < nk && st == PL_OK; ++i) {
    const int k = klist[i];
    e = pl::launch_gen_subsets(seed, 0, sample, nullptr, k, (int)n, idx, s);
    if (e != hipSuccess) {
      st = hip_fail(e, "autotune subsets");
      break;
    }
    Variant best{0, 0};
    float best_t = 0.f;
    for (int lanes : kLanes) {
      // the lane groups of 8..64 lanes compute bit-identical results, as do
      // the 128-lane paddings; the other kernels each have their own
      // summation order, so without `edges` only variants of the family that
      // k's default belongs to are offered and tuning never changes a result
      if (pl::variant_capacity(lanes) < k ||
          (!edges && family(lanes) != family(default_variant(k).lanes)))
        continue;
      for (int pad : kPads) {
        if (!pl::solve_variant_exists(lanes, pad)) continue;
        SolveJob j{G,   n, ld, idx, nullptr, k, sample, out, out + sample, out + 2 * sample,
                   (uint32_t *)(out + 3 * sample), -1,
                   {lanes, pad, use_f64(PL_FORM_TUNED, k, lanes)}};
        float t = 0.f;
        for (int rep = 0; rep < 3 && st == PL_OK; ++rep) {  // rep 0 warms up
          if ((e = hipEventRecord(e0, s)) != hipSuccess) st = hip_fail(e, "event");
          if (st == PL_OK) st = run_job(j, ws, s);
          if (st == PL_OK && (e = hipEventRecord(e1, s)) != hipSuccess) st = hip_fail(e, "event");
          if (st == PL_OK && (e = hipEventSynchronize(e1)) != hipSuccess)
            st = hip_fail(e, "autotune sync");
          float ms = 0.f;
          if (st == PL_OK && (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
            st = hip_fail(e, "event time");
          if (rep == 1 || (rep > 1 && ms < t)) t = ms;
        }
        if (st != PL_OK) break;
        if (sweep) {
          const int li = (int)(std::find(std::begin(kLanes), std::end(kLanes), lanes) - kLanes);
          const int pi = (int)(std::find(std::begin(kPads), std::end(kPads), pad) - kPads);
          sweep[i * PL_TUNE_VARIANTS + 3 * li + pi] = t;
        }
        if (best.lanes == 0 || t < best_t) {
          best = Variant{lanes, pad, false};
          best_t = t;
        }
      }
      if (st != PL_OK) break;
    }
    if (st != PL_OK) break;
    {
      std::lock_guard<std::mutex> lock(g_tune_mu);
      tune_init_locked();
      g_tune[k] = best;
    }
    if (best_ms) best_ms[i] = best_t;
  }
  if (st == PL_OK) {
    // bin edges: a size between two measured ones takes the winner of the
    // next measured size above it (whose lane count covers it); without
    // `edges` only when that winner belongs to the size's own kernel family,
    // else it keeps its default
    std::vector<int> ks(klist, klist + nk);
    std::sort(ks.begin(), ks.end());
    ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
    std::lock_guard<std::mutex> lock(g_tune_mu);
    size_t q = 0;
    for (int k = 1; k <= ks.back(); ++k) {
      while (ks[q] < k) ++q;
      if (ks[q] == k) continue;
      const Variant w = g_tune[ks[q]];
      g_tune[k] = (edges || family(w.lanes) == family(default_variant(k).lanes)) ? w
                                                                                 : default_variant(k);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (ws) (void)hipFree(ws);
  if (out) (void)hipFree(out);
  if (idx) (void)hipFree(idx);
  return st;
}

}  // namespace

extern "C" {

pl_status pl_autotune(const float *G, int64_t n, int64_t ld, const int32_t *klist, int nk,
                      int64_t sample, uint64_t seed, double *best_ms, void *stream) {
  return autotune(G, n, ld, klist, nk, sample, seed, false, best_ms, nullptr, stream);
}

pl_status pl_autotune_edges(const float *G, int64_t n, int64_t ld, const int32_t *klist, int nk,
                            int64_t sample, uint64_t seed, double *best_ms, double *sweep_ms,
                            void *stream) {
  return autotune(G, n, ld, klist, nk, sample, seed, true, best_ms, sweep_ms, stream);
}

pl_status pl_solve_alleig(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                          int64_t nsub, double *eig, double *det, int32_t *err, void *stream) {
  pl_status st = check_solve_args(G, n, ld, idx, k, nsub);
  if (st != PL_OK) return st;
  if (nsub > 0 && (!eig || !det || !err)) return fail(PL_ERR_ARG, "output pointers must be non-null");
  if (nsub == 0) return PL_OK;
  hipStream_t s = (hipStream_t)stream;
  Workspace w;
  st = w.acquire(nullptr, 0, pl::workspace_bytes(nsub, k), s);
  if (st != PL_OK) return st;
  const pl::WorkspaceParts wp = pl::split_workspace(w.p, nsub, k);
  float2 *tri = wp.tri;
  double4 *bounds = wp.bounds;
  const Variant v = job_variant(PL_FORM_TUNED, k);
  pl::SolveArgs a{};
  a.G = G;
  a.n = n;
  a.ld = ld;
  a.idx = idx;
  a.kuni = k;
  a.kcap = k;
  a.nslots = nsub;
  a.status = (uint32_t *)err;
  a.tri = tri;
  a.tstride = k;
  a.bounds = bounds;
  hipError_t e = pl::launch_solve(v.lanes, v.pad, pl::MODE_TRIDIAG, a, s, 0, v.f64);
  if (e != hipSuccess) return hip_fail(e, "all-eigenvalue solve launch");
  pl::BisectArgs b{};
  b.tri = tri;
  b.tstride = k;
  b.bounds = bounds;
  b.kuni = k;
  b.nsub = nsub;
  b.status = (uint32_t *)err;
  b.eig = eig;
  b.det = det;
  e = pl::launch_bisect(1, b, s);
  if (e != hipSuccess) return hip_fail(e, "all-eigenvalue bisection launch");
  return PL_OK;
}

pl_status pl_rip_stats(const float *lmin, const float *lmax, const float *delta,
                       const uint32_t *status, const int64_t *offsets, int k, int64_t nsub,
                       int kmax, double *out, void *stream) {
  if (!lmin || !lmax || !delta || !status || !out)
    return fail(PL_ERR_ARG, "pointers must be non-null");
  if (kmax < 1 || kmax > kMaxK) return fail(PL_ERR_ARG, "kmax = %d outside [1, %d]", kmax, kMaxK);
  if (!offsets && (k < 1 || k > kmax)) return fail(PL_ERR_ARG, "k = %d outside [1, kmax]", k);
  if (nsub < 0) return fail(PL_ERR_ARG, "nsub < 0");
  hipStream_t s = (hipStream_t)stream;
  unsigned *ext = nullptr;
  hipError_t e = hipMallocAsync((void **)&ext, sizeof(unsigned) * 3 * (kmax + 1), s);
  if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(rip scratch)");
  e = pl::launch_rip_stats(lmin, lmax, delta, status, offsets, offsets ? 0 : k, nsub, kmax, out,
                           ext, s);
  (void)hipFreeAsync(ext, s);
  if (e != hipSuccess) return hip_fail(e, "rip statistics launch");
  return PL_OK;
}

pl_status pl_gather(const float *G, int64_t n, int64_t ld, const int32_t *idx, int k,
                    int64_t nsub, float *out, int mode, int lanes, int pad, void *stream) {
  pl_status st = check_solve_args(G, n, ld, idx, k, nsub);
  if (st != PL_OK) return st;
  if (!out) return fail(PL_ERR_ARG, "out must be non-null");
  if (mode != 0 && mode != 1) return fail(PL_ERR_ARG, "gather mode must be 0 or 1");
  if (lanes == 0) lanes = group_for(k);
  if (!pl::solve_variant_exists(lanes, pad) || pl::variant_capacity(lanes) < k)
    return fail(PL_ERR_ARG, "no variant with %d lanes and pad %d for k = %d", lanes, pad, k);
  pl::SolveArgs a{};
  a.G = G;
  a.n = n;
  a.ld = ld;
  a.idx = idx;
  a.kuni = k;
  a.kcap = k;
  a.nslots = nsub;
  a.gs_out = out;
  hipError_t e = pl::launch_solve(lanes, pad, mode == 0 ? pl::MODE_DUMP : pl::MODE_CHECKSUM, a,
                                  (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gather launch");
  return PL_OK;
}

pl_status pl_gen_subsets(uint64_t seed, int64_t s0, int64_t nsub, int k, int n, int32_t *idx,
                         void *stream) {
  if (!idx) return fail(PL_ERR_ARG, "idx must be non-null");
  if (n < 1 || k < 1 || k > n) return fail(PL_ERR_ARG, "need 1 <= k <= n (k=%d, n=%d)", k, n);
  if (nsub < 0 || s0 < 0) return fail(PL_ERR_ARG, "negative subset range");
  hipError_t e = pl::launch_gen_subsets(seed, s0, nsub, nullptr, k, n, idx, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gen_subsets launch");
  return PL_OK;
}

pl_status pl_gen_subsets_csr(uint64_t seed, int64_t s0, int64_t nsub, const int64_t *offsets,
                             int n, int32_t *idx, void *stream) {
  if (!idx || !offsets) return fail(PL_ERR_ARG, "idx and offsets must be non-null");
  if (n < 1 || nsub < 0 || s0 < 0) return fail(PL_ERR_ARG, "bad n or subset range");
  hipError_t e = pl::launch_gen_subsets(seed, s0, nsub, offsets, 0, n, idx, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gen_subsets_csr launch");
  return PL_OK;
}

pl_status pl_gen_ks(uint64_t seed, int64_t s0, int64_t nsub, const int32_t *klist, int nk,
                    int32_t *ks, void *stream) {
  if (!klist || !ks) return fail(PL_ERR_ARG, "klist and ks must be non-null");
  if (nk < 1 || nk > 32) return fail(PL_ERR_ARG, "nk = %d outside [1, 32]", nk);
  if (nsub < 0 || s0 < 0) return fail(PL_ERR_ARG, "negative subset range");
  hipError_t e = pl::launch_gen_ks(seed, s0, nsub, klist, nk, ks, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gen_ks launch");
  return PL_OK;
}

pl_status pl_gen_phi(int kind, uint64_t seed, int m, int n, float *phi, void *stream) {
  if (!phi) return fail(PL_ERR_ARG, "phi must be non-null");
  if (m < 1 || n < 1) return fail(PL_ERR_ARG, "m and n must be positive");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (kind == 0) {
    e = pl::launch_gen_phi_gauss(seed, m, n, phi, s);
  } else if (kind == 1) {
    e = pl::launch_gen_phi_bernoulli(seed, m, n, phi, s);
  } else if (kind == 2) {
    if (m > n) return fail(PL_ERR_ARG, "DCT rows: m = %d > n = %d", m, n);
    // row set: Floyd sample of m rows of [0,n) from stream (seed, TAG_DCT, 0)
    std::vector<int32_t> rows((size_t)m);
    pl::Stream st(seed, pl::TAG_DCT, 0u);
    int cnt = 0;
    for (int j = n - m; j < n; ++j) {
      const int32_t t = (int32_t)st.bounded((uint32_t)j + 1u);
      bool seen = false;
      for (int q = 0; q < cnt && !seen; ++q) seen = rows[(size_t)q] == t;
      rows[(size_t)cnt++] = seen ? j : t;
    }
    int32_t *drows = nullptr;
    e = hipMalloc(&drows, sizeof(int32_t) * (size_t)m);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(dct rows)");
    e = hipMemcpyAsync(drows, rows.data(), sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = pl::launch_gen_phi_dct(drows, m, n, phi, s);
    hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    (void)hipFree(drows);
  } else {
    return fail(PL_ERR_ARG, "unknown phi kind %d", kind);
  }
  if (e != hipSuccess) return hip_fail(e, "gen_phi");
  return PL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Drop-in host entry points (ParaLanczos.h)
// ---------------------------------------------------------------------------
namespace {

std::mutex g_legacy_mu;

bool quiet() {
  const char *q = getenv("PARALANCZOS_QUIET");
  return q && *q && strcmp(q, "0") != 0;
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  hipError_t alloc(size_t count) { return hipMalloc(&p, sizeof(T) * (count ? count : 1)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// Uploads mat (FP64 -> FP32) and the index sets, runs the all-eigenvalue solve
// and copies back n_eig eigenvalues, n_det determinants and n_err status words.
// Returns 0 on success, -1 (with a message on stderr) on failure.
int legacy_solve(const double *mat, int matsize, const int *subIndex, int L, int subn,
                 double *eig_out, int64_t n_eig, double *det_out, int64_t n_det, int *err_out,
                 int64_t n_err) {
  if (subn < 1 || subn > kMaxK) {
    fprintf(stderr, " subset size %d is not supported by this build (1..%d) \n", subn, kMaxK);
    return -1;
  }
  if (L < subn || L % subn != 0 || matsize < 1) {
    fprintf(stderr, " index length %d is not a positive multiple of the subset size %d \n", L,
            subn);
    return -1;
  }
  for (int i = 0; i < L; ++i)
    if (subIndex[i] < 0 || subIndex[i] >= matsize) {
      fprintf(stderr, " subIndex[%d] = %d is outside [0, %d) \n", i, subIndex[i], matsize);
      return -1;
    }
  const int64_t nsub = L / subn;
  const size_t nn = (size_t)matsize * (size_t)matsize;
  std::vector<float> g32(nn);
  for (size_t i = 0; i < nn; ++i) g32[i] = (float)mat[i];
  DevBuf<float> dG;
  DevBuf<int32_t> dIdx, dErr;
  DevBuf<double> dEig, dDet;
  hipError_t e = dG.alloc(nn);
  if (e == hipSuccess) e = dIdx.alloc((size_t)L);
  if (e == hipSuccess) e = dEig.alloc((size_t)L);
  if (e == hipSuccess) e = dDet.alloc((size_t)nsub);
  if (e == hipSuccess) e = dErr.alloc((size_t)nsub);
  if (e == hipSuccess) e = hipMemcpy(dG.p, g32.data(), sizeof(float) * nn, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dIdx.p, subIndex, sizeof(int32_t) * (size_t)L, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fprintf(stderr, " ParaLanczos: %s \n", hipGetErrorString(e));
    return -1;
  }
  pl_status st =
      pl_solve_alleig(dG.p, matsize, matsize, dIdx.p, subn, nsub, dEig.p, dDet.p, dErr.p, nullptr);
  if (st != PL_OK) {
    fprintf(stderr, " ParaLanczos: %s \n", pl_last_error());
    return -1;
  }
  e = hipStreamSynchronize(nullptr);
  if (e == hipSuccess && n_eig > 0)
    e = hipMemcpy(eig_out, dEig.p, sizeof(double) * (size_t)n_eig, hipMemcpyDeviceToHost);
  if (e == hipSuccess && n_det > 0)
    e = hipMemcpy(det_out, dDet.p, sizeof(double) * (size_t)n_det, hipMemcpyDeviceToHost);
  if (e == hipSuccess && n_err > 0)
    e = hipMemcpy(err_out, dErr.p, sizeof(int32_t) * (size_t)n_err, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    fprintf(stderr, " ParaLanczos: %s \n", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

void print_eigs(const double *eig, int count, const char *word) {
  if (quiet()) return;
  for (int i = 0; i < count; ++i) printf("%d %s: %g \n", i, word, eig[i]);
  fflush(stdout);
}

}  // namespace

extern "C" {

void print_tridiagonal(double *alpha, double *beta, int n) {
  printf(" [ ");
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < n; ++c) {
      double x = 0.0;
      if (r == c) x = alpha[c];
      else if (c == r - 1) x = beta[c + 1];
      else if (c == r + 1) x = beta[c];
      printf(" %g ", x);
      if (c != n - 1) printf(", ");
    }
    printf(r != n - 1 ? "; \n" : "] \n");
  }
  fflush(stdout);
}

void calculate_eigs(double *mat, int n, double *Eigenvalues, double *Det) {
  if (n % 4 != 0 || n > 128 || n < 1) {
    fprintf(stderr, " 'int n' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  std::vector<int> ident((size_t)n);
  for (int i = 0; i < n; ++i) ident[(size_t)i] = i;
  if (legacy_solve(mat, n, ident.data(), n, n, Eigenvalues, n, Det, 1, nullptr, 0) != 0) return;
  print_eigs(Eigenvalues, n, "eigevalue");
}

void calculate_subeigs(double *mat, int n, double *Eigenvalues, double *Det, int *subIndex,
                       int subn) {
  if (subn % 4 != 0 || n > 128 || subn < 1) {
    fprintf(stderr, " 'int n' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  if (legacy_solve(mat, n, subIndex, subn, subn, Eigenvalues, subn, Det, 1, nullptr, 0) != 0)
    return;
  print_eigs(Eigenvalues, subn, "eigenvalue");
}

void calculate_subeigs2(double *mat, int n, int matsize, double *Eigenvalues, double *Det,
                        int *subIndex, int subn) {
  if (subn % 4 != 0 || n > 128 || subn < 1) {
    fprintf(stderr, " 'int n' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  if (!quiet()) {
    printf("welcome to calculate subeigs2...\n");
    fflush(stdout);
  }
  if (legacy_solve(mat, matsize, subIndex, n, subn, Eigenvalues, n, Det, 1, nullptr, 0) != 0)
    return;
  print_eigs(Eigenvalues, n, "eigenvalue");
}

void calculate_subeigs3(double *mat, int subIndexLen, int matsize, double *Eigenvalues,
                        double *Det, int *subIndex, int subn) {
  if (subn % 4 != 0 || subn > 128 || subn < 1) {
    fprintf(stderr, " 'int subn' must be divisible by 4, and less than or equal to 128! \n");
    return;
  }
  if (subIndexLen % 32 != 0) {
    fprintf(stderr, " 'subIndexLen' must be divisible by 32! \n");
    return;
  }
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  if (!quiet()) {
    printf("welcome to calculate subeigs2...\n");
    fflush(stdout);
  }
  if (legacy_solve(mat, matsize, subIndex, subIndexLen, subn, Eigenvalues, subIndexLen, Det,
                   subIndexLen / subn, nullptr, 0) != 0)
    return;
  print_eigs(Eigenvalues, subIndexLen, "eigenvalue");
}

int pl_multi_subsvd(const double *g, int M, const int *subIndex, int L, int minor, double *eigs,
                    double *det, int *err) {
  if (!g || !subIndex || !eigs || !det || !err || M < 1 || minor < 1 || L < minor ||
      L % minor != 0)
    return -1;
  std::lock_guard<std::mutex> lock(g_legacy_mu);
  return legacy_solve(g, M, subIndex, L, minor, eigs, L, det, L / minor, err, L / minor);
}

}  // extern "C"
