// tgrav: C ABI of the MI355X direct-sum gravity field (see include/tgrav.h).
//
// Host side: argument checking, host<->device staging for TGRAV_HOST pointers, source
// packing, launch-shape selection (kernel variant + source chunking) and the chunk
// reduction.  Device code lives in kernels.hpp (grid kernels) and batch.hpp
// (persistent batch kernel).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tgrav.h"
#include "kernels.hpp"
#include "batch.hpp"
#include "bidir.hpp"
#include "grid.hpp"
#include "flash.hpp"
#include "pm.hpp"

#define TGRAV_VERSION 1

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHECK(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(TGRAV_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                  __FILE__, __LINE__);                                                   \
  } while (0)

#define LAUNCHCHECK(what)                                                                \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess)                                                                \
      return fail(TGRAV_EHIP, "launch of %s failed: %s", what, hipGetErrorString(e_));    \
  } while (0)

// ---- fixed scratch of the allocation-free entry point (tgrav_eval_static) ----
// One array per GPU the library runs on, held until the process ends.  Host code never passes
// `scratch` itself (on the host it names a stand-in, not device memory): it passes the device
// address hipGetSymbolAddress gave for the current device, kept per device in g_scratch_ptr.
#define TG_STATIC_SCRATCH_BYTES (64ull << 20)
#define TG_MAX_DEVICES 64
#ifndef TG_COL_WAVES
#define TG_COL_WAVES 4   // default waves per column workgroup (tgrav_set_column_waves)
#endif
static __device__ __align__(256) unsigned char scratch[TG_STATIC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[TG_MAX_DEVICES] = {};

// Once per device, with that device current (called by tgrav_create).
static int scratch_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= TG_MAX_DEVICES)
    return fail(TGRAV_ELIMIT, "device %d beyond the %d the library supports", device, TG_MAX_DEVICES);
  if (g_scratch_ptr[device]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(TGRAV_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return 0;
}

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return 0;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return fail(TGRAV_EHIP, "hipFree failed: %s", hipGetErrorString(e));
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&p, want) != hipSuccess) {
      p = nullptr;
      return fail(TGRAV_ENOMEM, "hipMalloc of %zu bytes failed", want);
    }
    cap = want;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const { return (T*)p; }
};

// ---------------------------------------------------------------- variants ------
typedef void (*launch_fn)(dim3, dim3, hipStream_t, const tg::Tgt&, const double4*, const double*,
                          int64_t, int64_t, const double*, const tg::Out&);
typedef void (*batch_fn)(int, hipStream_t, const tg::BProb*, const tg::BItem*, int, unsigned int*,
                         const double4*, const double*, const unsigned long long*, const int*,
                         double*);
typedef const void* (*kernel_ptr_fn)();

template <int K, int NT, bool MIXED>
void launch_batch(int grid, hipStream_t s, const tg::BProb* p, const tg::BItem* it, int n,
                  unsigned int* ctr, const double4* s4, const double* se2,
                  const unsigned long long* st, const int* colflag, double* part) {
  tg::k_batch_eval<K, NT, MIXED><<<grid, NT, 0, s>>>(p, it, n, ctr, s4, se2, st, colflag, part);
}
template <int K, int NT, bool MIXED>
const void* batch_kernel() {
  return (const void*)tg::k_batch_eval<K, NT, MIXED>;
}

template <int K, int NT>
void launch_fp64(dim3 g, dim3 b, hipStream_t s, const tg::Tgt& T, const double4* s4,
                 const double* se2, int64_t ns, int64_t chunk, const double* e2m, const tg::Out& O) {
  tg::k_eval_fp64<K, NT><<<g, b, 0, s>>>(T, s4, se2, ns, chunk, e2m, O);
}
template <int K, int NT>
void launch_mixed(dim3 g, dim3 b, hipStream_t s, const tg::Tgt& T, const double4* s4,
                  const double* se2, int64_t ns, int64_t chunk, const double* e2m, const tg::Out& O) {
  tg::k_eval_mixed<K, NT><<<g, b, 0, s>>>(T, s4, se2, ns, chunk, e2m, O);
}

struct Variant {
  int k, nt, ts;
  launch_fn fn;
  batch_fn bfn;
  kernel_ptr_fn bkernel;
};

#define FP64V(K, NT) {K, NT, NT, launch_fp64<K, NT>, launch_batch<K, NT, false>, batch_kernel<K, NT, false>}
#define MIXV(K, NT) {K, NT, NT, launch_mixed<K, NT>, launch_batch<K, NT, true>, batch_kernel<K, NT, true>}
const Variant kFp64Variants[] = {FP64V(1, 256), FP64V(2, 256), FP64V(4, 256),
                                 FP64V(2, 128), FP64V(4, 128), FP64V(8, 64)};
const Variant kMixedVariants[] = {MIXV(2, 256), MIXV(4, 256), MIXV(8, 256),
                                  MIXV(4, 128), MIXV(8, 128), MIXV(16, 64)};
#undef FP64V
#undef MIXV
typedef void (*bidir_fn)(dim3, dim3, hipStream_t, const tg::Cells&, const double4*, const double*,
                         int64_t, int64_t, const double*, const tg::Out&, double*);
template <int K, int NT, bool MIXED>
void launch_bidir(dim3 g, dim3 b, hipStream_t s, const tg::Cells& C, const double4* p4,
                  const double* pe2, int64_t np, int64_t chunk, const double* e2m,
                  const tg::Out& CO, double* ppart) {
  tg::k_bidir<K, NT, MIXED><<<g, b, 0, s>>>(C, p4, pe2, np, chunk, e2m, CO, ppart);
}
struct BiVariant {
  int k, nt;
  bidir_fn fn;
};
const BiVariant kBiVariants[2][4] = {
    {{2, 256, launch_bidir<2, 256, false>}, {4, 256, launch_bidir<4, 256, false>},
     {4, 128, launch_bidir<4, 128, false>}, {8, 128, launch_bidir<8, 128, false>}},
    {{2, 256, launch_bidir<2, 256, true>}, {4, 256, launch_bidir<4, 256, true>},
     {4, 128, launch_bidir<4, 128, true>}, {8, 128, launch_bidir<8, 128, true>}}};
const int kNumBi = 4;

const int kNumFp64 = sizeof(kFp64Variants) / sizeof(kFp64Variants[0]);
const int kNumMixed = sizeof(kMixedVariants) / sizeof(kMixedVariants[0]);
const int kDefaultFp64 = 1;
const int kDefaultMixed = 1;

const Variant* variant_of(int mode, int v) {
  if (mode == TGRAV_PLAIN_FP64) return (v >= 0 && v < kNumFp64) ? &kFp64Variants[v] : nullptr;
  if (mode == TGRAV_MIXED) return (v >= 0 && v < kNumMixed) ? &kMixedVariants[v] : nullptr;
  return nullptr;
}

// ------------------------------------------------------------- tune cache ------
// The autotuner's JSON cache (torch_amd/autotune.py writes it): a flat object whose keys are
// "<arch>/<CUs>cu|<kind>|<mode>|<size buckets>" and whose values are objects of numbers
// (variant, nsplit, item, wg_per_cu, ...).  Size buckets are round(log2(n)).
struct TuneEntry {
  int variant = -1, nsplit = -1, wg_per_cu = -1;
  double item = -1.0, col_scale = -1.0, gs_scale = -1.0;
};

struct JsonIn {
  const char* p;
  const char* e;
  void ws() {
    while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
  }
  bool eat(char ch) {
    ws();
    if (p < e && *p == ch) { ++p; return true; }
    return false;
  }
  bool str(std::string& out) {
    ws();
    if (p >= e || *p != '"') return false;
    ++p;
    out.clear();
    while (p < e && *p != '"') {
      if (*p == '\\' && p + 1 < e) ++p;
      out.push_back(*p++);
    }
    if (p >= e) return false;
    ++p;
    return true;
  }
  bool num(double& v) {
    ws();
    char* end = nullptr;
    v = strtod(p, &end);
    if (end == p || end > e) return false;
    p = end;
    return true;
  }
  bool skip_value() {   // a number, string, literal or nested object/array
    ws();
    if (p >= e) return false;
    if (*p == '"') { std::string t; return str(t); }
    if (*p == '{' || *p == '[') {
      int depth = 0;
      bool in_str = false;
      for (; p < e; ++p) {
        if (in_str) { if (*p == '\\') ++p; else if (*p == '"') in_str = false; continue; }
        if (*p == '"') in_str = true;
        else if (*p == '{' || *p == '[') ++depth;
        else if (*p == '}' || *p == ']') { if (--depth == 0) { ++p; return true; } }
      }
      return false;
    }
    const char* start = p;   // true / false / null
    while (p < e && *p != ',' && *p != '}' && *p != ']' && *p != ' ' && *p != '\n') ++p;
    return p > start;
  }
};

// Parses the cache text; returns the number of entries or -1 on malformed input.
static int parse_tune_cache(const std::string& text, std::map<std::string, TuneEntry>* out) {
  JsonIn in{text.data(), text.data() + text.size()};
  if (!in.eat('{')) return -1;
  int n = 0;
  if (in.eat('}')) return 0;
  do {
    std::string key;
    if (!in.str(key) || !in.eat(':') || !in.eat('{')) return -1;
    TuneEntry t;
    if (!in.eat('}')) {
      do {
        std::string f;
        if (!in.str(f) || !in.eat(':')) return -1;
        in.ws();
        if (in.p < in.e && (*in.p == '-' || (*in.p >= '0' && *in.p <= '9'))) {
          double v;
          if (!in.num(v)) return -1;
          if (f == "variant") t.variant = (int)v;
          else if (f == "nsplit") t.nsplit = (int)v;
          else if (f == "wg_per_cu") t.wg_per_cu = (int)v;
          else if (f == "item") t.item = v;
          else if (f == "col_scale") t.col_scale = v;
          else if (f == "gs_scale") t.gs_scale = v;
        } else if (!in.skip_value()) {
          return -1;
        }
      } while (in.eat(','));
      if (!in.eat('}')) return -1;
    }
    if (out) (*out)[key] = t;
    ++n;
  } while (in.eat(','));
  if (!in.eat('}')) return -1;
  return n;
}

static int read_file(const char* path, std::string* text) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  char buf[65536];
  size_t k;
  text->clear();
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) text->append(buf, k);
  fclose(f);
  return 0;
}

static int size_bucket(double n) { return (int)lround(log2(std::max(n, 1.0))); }

}  // namespace

struct tgrav_ctx {
  int device = 0;
  int num_cu = 256;
  std::string devkey;                          // "<arch>/<CUs>cu", the tune-cache key prefix
  std::map<std::string, TuneEntry> tune;       // loaded tune cache (may be empty)
  bool user_cfg[2] = {false, false};           // tgrav_set_config was called: it wins
  bool user_batch = false;                     // tgrav_set_batch_config was called
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  double G = 4.300917270e-3;
  int validate = 1;
  int variant[2] = {kDefaultFp64, kDefaultMixed};
  int nsplit[2] = {0, 0};
  int block_kernel = 1;      // 8^3 block targets use the column kernel (k_eval_grid)
  int col_waves = TG_COL_WAVES;   // waves per column workgroup (k_eval_grid / cols, batch: 1, 2, 4)
  int bvariant[2] = {1, 1};  // fused bidirectional kernel: variant and particle chunks
  int bnsplit[2] = {0, 0};
  // softening kernel: TGRAV_SOFT_* (FLASH kernels use the global radius rsoft), and the
  // periodic image sum of the FLASH kernels
  int soft = TGRAV_SOFT_PLUMMER;
  double rsoft = 0.0;
  int soft_quirks = 0;
  int periodic = 0;           // 1: periodic images (2: always the per-target image walk)
  double box[3] = {0.0, 0.0, 0.0};
  int nrep = 6;
  double batch_item = 0.0;   // target interactions per batch work item (0 = automatic)
  int batch_wg_per_cu = 0;   // persistent workgroups per CU (0 = occupancy limit)
  double batch_scale[2] = {1.0, 1.0};   // column / block-source work lists: items of this many
  bool user_scales = false;             //   times the item size; tgrav_set_batch_scales was called
  // packed sources; stat: their statistics (tg::ST_*, read by the kernels; ST_BAD = NaN flag)
  int64_t ns = -1;
  DevBuf s4, se2, stat;
  // sources given as 8^3 blocks: descriptors {ox, oy, oz, dx} and eps^2 per block
  int64_t src_nblk = 0;      // > 0: the current source set is block-structured
  DevBuf sbe2, sax, sms;     // source blocks: eps^2, cell-centre axes (24 per block), cell masses
  int64_t src_cols = 0;      // > 0: explicit sources that may be 8^3 blocks (scolflag says)
  DevBuf scolflag;
  // staging for host pointers and scratch; flag[0] = NaN flag of the targets
  DevBuf in[12], outb[8], vel[3], part, flag;
  DevBuf colflag;            // explicit targets: 1 if they have 8^3 column structure
  DevBuf bp4, bpe2, bstat, bppart, bpgrp;   // fused passes: packed particles, partials
  DevBuf pm[16], pm_acc[3], pm_stat;  // particle-mesh pieces: staging, fixed-point sums, stat
  DevBuf rstat;              // range statistics of the mixed mode's check (2 x tg::ST_WORDS)
  DevBuf pkey, pidx, ptmp, pbuf;   // periodic pairs: sort keys and indices (2 x each), sort scratch,
                                   //   per-pair image sums
  unsigned long long* host_flag = nullptr;   // pinned, kHostWords words
  // instrumentation: HIP events around each evaluation kernel (tgrav_set_timing)
  int timing = 0;
  std::vector<hipEvent_t> ev;                // 2 per slot
  std::vector<int64_t> ev_nt, ev_ns;
  int ev_used = 0;
};

static const TuneEntry* tune_lookup(const tgrav_ctx* c, const std::string& key) {
  if (c->tune.empty()) return nullptr;
  auto it = c->tune.find(key);
  return it == c->tune.end() ? nullptr : &it->second;
}

static const char* mode_name(int mode) { return mode == TGRAV_MIXED ? "mixed" : "fp64"; }

// Tuned source split of a column kernel ("blockfwd" / "blocksrc"), or -1.  The key's sizes are
// the explicit point count (the sources of "blockfwd", the targets of "blocksrc") and the
// number of leaf blocks, as the autotuner writes them.
static int tuned_block_split(const tgrav_ctx* c, int mode, const char* dir, int64_t ns, int64_t nblk,
                             int waves = 1) {
  if (c->user_cfg[mode] || c->tune.empty()) return -1;
  char key[256], d[64];
  if (waves > 1) snprintf(d, sizeof d, "%s/w%d", dir, waves);   // a split tuned for this workgroup
  else snprintf(d, sizeof d, "%s", dir);
  snprintf(key, sizeof key, "%s|%s|%s|s%d|b%d", c->devkey.c_str(), d, mode_name(mode),
           size_bucket((double)ns), size_bucket((double)nblk));
  const TuneEntry* t = tune_lookup(c, key);
  return t ? t->nsplit : -1;
}

static const int kTimingSlots = 256;
static const int kHostWords = 16;   // pinned host words: flags and two sets of range statistics

template <class F>
static int timed_launch_on(tgrav_ctx* c, hipStream_t s, int64_t nt, int64_t ns, F launch) {
  if (!c->timing || c->ev_used >= kTimingSlots) return launch();
// Records an event pair around `launch` when timing is on (slots beyond the pool are not timed).
  if (c->ev.empty()) {
    c->ev.resize(2 * kTimingSlots, nullptr);
    for (auto& e : c->ev) HIPCHECK(hipEventCreate(&e));
    c->ev_nt.resize(kTimingSlots);
    c->ev_ns.resize(kTimingSlots);
  }
  const int k = c->ev_used++;
  HIPCHECK(hipEventRecord(c->ev[2 * k], s));
  int rc = launch();
  if (rc) return rc;
  HIPCHECK(hipEventRecord(c->ev[2 * k + 1], s));
  c->ev_nt[k] = nt;
  c->ev_ns[k] = ns;
  return 0;
}

template <class F>
static int timed_launch(tgrav_ctx* c, int64_t nt, int64_t ns, F launch) {
  return timed_launch_on(c, c->stream, nt, ns, launch);
}

// ------------------------------------------------------------------ helpers ------
static int set_device(tgrav_ctx* c) {
  HIPCHECK(hipSetDevice(c->device));
  return 0;
}

static int check_mode(int mode) {
  if (mode != TGRAV_PLAIN_FP64 && mode != TGRAV_MIXED)
    return fail(TGRAV_EINVAL, "mode must be TGRAV_PLAIN_FP64 (0) or TGRAV_MIXED (1), got %d", mode);
  return 0;
}

static int check_kind(int kind) {
  if (kind != TGRAV_HOST && kind != TGRAV_DEVICE)
    return fail(TGRAV_EINVAL, "ptr_kind must be TGRAV_HOST (0) or TGRAV_DEVICE (1), got %d", kind);
  return 0;
}

static bool host_finite(int64_t n, const double* a) {
  if (!a) return true;
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(a[i])) return false;
  return true;
}

// Copies up to 5 host arrays of n doubles into staging buffers in[first..]; NULL stays NULL.
static int stage_in(tgrav_ctx* c, int first, int64_t n, const double** arrs, int cnt) {
  for (int q = 0; q < cnt; ++q) {
    if (!arrs[q]) continue;
    int rc = c->in[first + q].ensure((size_t)n * 8);
    if (rc) return rc;
    HIPCHECK(hipMemcpyAsync(c->in[first + q].p, arrs[q], (size_t)n * 8, hipMemcpyHostToDevice,
                            c->stream));
    arrs[q] = c->in[first + q].as<double>();
  }
  return 0;
}

// Workgroups for n lanes of bs (at least one: every kernel launched this way checks its index
// against n, so an empty input is a no-op launch rather than an invalid configuration).
static unsigned grid1(int64_t n, int bs) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

// Sums nsplit chunk partials (4 x pstride doubles each) into the outputs of n targets.  Few
// targets with many chunks take the wide kernel (16 waves per 64 targets); otherwise one
// thread per target walks the chunks.
static int reduce_chunks(tgrav_ctx* c, const double* part, int64_t pstride, int nsplit, int64_t n,
                         const tg::Out& fin) {
  if (n < 256 * 1024 && nsplit >= 16) {
    tg::k_reduce_chunks_wide<<<grid1(n, 64), 1024, 0, c->stream>>>(part, pstride, nsplit, n, fin);
    LAUNCHCHECK("k_reduce_chunks_wide");
  } else {
    tg::k_reduce_chunks<<<grid1(n, 256), 256, 0, c->stream>>>(part, pstride, nsplit, n, fin);
    LAUNCHCHECK("k_reduce_chunks");
  }
  return 0;
}

// Reads a device flag word written by k_pack_sources / k_check_finite (synchronous).
static int read_flag(tgrav_ctx* c, const void* dflag, bool* bad) {
  HIPCHECK(hipMemcpyAsync(c->host_flag, dflag, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  *bad = *c->host_flag != 0ULL;
  return 0;
}

// ---- range of the mixed mode (only with validation on) ----
// The mixed kernels hold displacements, r^2 and m / r^3 in FP32.  With S and T the statistics of
// the two point sets (tg::ST_*), every r^2 is below 3 (R_S + R_T)^2 + eps_S^2 + eps_T^2 and every
// weight m y^3 below M / lo^(3/2), lo = the smallest eps_S^2 + eps_T^2 (0: unsoftened sets, whose
// pairs may come arbitrarily close whatever the precision, are not bounded here).
#define TG_MIXED_R2_MAX 1e36
#define TG_MIXED_W_MAX 1e30

static int mixed_range_ok(const double* S, const double* T, const char* what) {
  const double d = S[tg::ST_RMAX] + T[tg::ST_RMAX];
  const double r2 = 3.0 * d * d + S[tg::ST_E2MAX] + T[tg::ST_E2MAX];
  if (!(r2 <= TG_MIXED_R2_MAX))
    return fail(TGRAV_ELIMIT, "mixed precision: %s span |coordinates| up to %.3g, so r^2 may reach %.3g, "
                "beyond the FP32 range of the mixed kernels (%.0e); use TGRAV_PLAIN_FP64 or larger units",
                what, std::max(S[tg::ST_RMAX], T[tg::ST_RMAX]), r2, TG_MIXED_R2_MAX);
  const double lo = S[tg::ST_E2MIN] + T[tg::ST_E2MIN];
  if (lo > 0.0 && !(S[tg::ST_MMAX] <= TG_MIXED_W_MAX * lo * sqrt(lo)))
    return fail(TGRAV_ELIMIT, "mixed precision: %s have |m| up to %.3g with softening^2 down to %.3g, so "
                "m / r^3 may exceed the FP32 range of the mixed kernels; use TGRAV_PLAIN_FP64 or other units",
                what, S[tg::ST_MMAX], lo);
  return 0;
}

// Statistics (tg::ST_*) of a device point set (dx == null) or block set (x, y, z corners, dx cell
// widths) with nm masses, into out[slot * ST_WORDS ...] of the pinned host words (synchronous).
static int range_stats(tgrav_ctx* c, int slot, int64_t n, const double* x, const double* y,
                       const double* z, const double* dx, const double* e, int64_t nm,
                       const double* m) {
  int rc = c->rstat.ensure(2 * tg::ST_WORDS * sizeof(unsigned long long));
  if (rc) return rc;
  unsigned long long* d = c->rstat.as<unsigned long long>() + slot * tg::ST_WORDS;
  tg::k_reset_stats<<<1, 64, 0, c->stream>>>(d);
  LAUNCHCHECK("k_reset_stats");
  if (std::max(n, nm) > 0) {
    tg::k_range<<<grid1(std::max(n, nm), 256), 256, 0, c->stream>>>(n, x, y, z, dx, e, nm, m, d);
    LAUNCHCHECK("k_range");
  }
  HIPCHECK(hipMemcpyAsync(c->host_flag + slot * tg::ST_WORDS, d, tg::ST_WORDS * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

static const double* host_stats(const tgrav_ctx* c, int slot) {
  return reinterpret_cast<const double*>(c->host_flag + slot * tg::ST_WORDS);
}

// Mixed-mode range check of an evaluation: the current source set (its device statistics) against
// targets given as points or blocks (device pointers).
static int eval_range_check(tgrav_ctx* c, const tg::Tgt& T, int64_t nin) {
  int rc;
  if (T.blocks) rc = range_stats(c, 1, nin, T.ox, T.oy, T.oz, T.bdx, T.beps, 0, nullptr);
  else rc = range_stats(c, 1, nin, T.x, T.y, T.z, nullptr, T.eps, 0, nullptr);
  if (rc) return rc;
  if (c->ns > 0) {
    HIPCHECK(hipMemcpyAsync(c->host_flag, c->stat.p, tg::ST_WORDS * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if ((rc = mixed_range_ok(host_stats(c, 0), host_stats(c, 1), "the sources and targets"))) return rc;
  }
  return 0;
}

// NaN/Inf check of device target arrays (synchronous; only with validation on).
static int device_check_finite(tgrav_ctx* c, int64_t n, const double* const* chk, int nchk,
                               bool* bad) {
  *bad = false;
  if (n <= 0) return 0;
  int rc = c->flag.ensure(sizeof(unsigned long long));
  if (rc) return rc;
  const double* cs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int q = 0; q < nchk && q < 5; ++q) cs[q] = chk[q];
  HIPCHECK(hipMemsetAsync(c->flag.p, 0, sizeof(unsigned long long), c->stream));
  tg::k_check_finite<<<grid1(n, 256), 256, 0, c->stream>>>(n, cs[0], cs[1], cs[2], cs[3], cs[4],
                                                          c->flag.as<unsigned long long>());
  LAUNCHCHECK("k_check_finite");
  return read_flag(c, c->flag.p, bad);
}

// Launch shape: the variant's target tile and a source split that gives about
// 4 workgroups per CU when the target set alone cannot fill the chip.
static void choose_split(const tgrav_ctx* c, int mode, const Variant* v, int64_t nt, int64_t ns,
                         int64_t* ntiles, int* nsplit, int64_t* chunk, int force = -1) {
  const int64_t tile = (int64_t)v->k * v->nt;
  *ntiles = (nt + tile - 1) / tile;
  int want = force >= 0 ? force : c->nsplit[mode];
  if (want <= 0) {
    // About 8 rounds of workgroups at ~3 resident waves per SIMD, at least 4 source tiles a
    // chunk (1 when that would leave fewer workgroups than CUs: small problems, where the
    // partial sums are few), and at most 12 x 2^20 target-chunks of partial sums (32 B each).
    const int64_t target_wg = 8LL * 12 * c->num_cu / (v->nt / 64);
    int64_t s = (target_wg + *ntiles - 1) / *ntiles;
    int64_t maxs = (ns + 4LL * v->ts - 1) / (4LL * v->ts);
    if (*ntiles * maxs < c->num_cu) maxs = (ns + v->ts - 1) / v->ts;
    const int64_t maxp = std::max<int64_t>(1, (12LL << 20) / std::max<int64_t>(nt, 1));
    s = std::min<int64_t>(s, std::min<int64_t>(maxs, maxp));
    want = (int)(s < 1 ? 1 : s);
  }
  if (want > 65535) want = 65535;
  int64_t ch = (ns + want - 1) / want;
  ch = (ch + v->ts - 1) / v->ts * v->ts;
  if (ch < v->ts) ch = v->ts;
  *chunk = ch;
  *nsplit = (int)((ns + ch - 1) / ch);
  if (*nsplit < 1) *nsplit = 1;
}

// 8^3 block targets: one workgroup of W = col_waves waves per leaf block (k_eval_grid; the waves
// split the workgroup's sources), sources split further over blockIdx.y.
static void grid_shape(const tgrav_ctx* c, int mode, int64_t ns, int64_t nblk, int* nsplit,
                       int64_t* chunk) {
  const int W = c->col_waves;
  int want = c->nsplit[mode];
  const int tuned = tuned_block_split(c, mode, "blockfwd", ns, nblk, W);
  if (tuned > 0) want = tuned;
  if (want <= 0) {
    // ~3 waves resident per SIMD: aim at ~8 rounds of waves so the last round's imbalance is
    // small, with at most 12 chunks (each adds 32 B of partial sums per target) and at least 4
    // source tiles per wave and chunk -- unless that leaves fewer waves than CUs (few blocks, few
    // sources): then down to one source tile per wave and chunk.
    const int64_t target = 8LL * 12 * c->num_cu;
    const int64_t per = (target + nblk * W - 1) / (nblk * W);
    int64_t w = std::min<int64_t>(per, std::min<int64_t>(12, std::max<int64_t>(1, ns / (256 * W))));
    if (nblk * W * w < c->num_cu) w = std::min<int64_t>(per, std::max<int64_t>(1, ns / (64 * W)));
    want = (int)std::min<int64_t>(w, 65535);
  }
  if (want > 65535) want = 65535;
  int64_t ch = (ns + want - 1) / want;
  ch = (ch + 63) / 64 * 64;
  *chunk = ch;
  *nsplit = (int)((ns + ch - 1) / ch);
}

// The column kernels for the context's waves per workgroup: k_eval_grid (block targets) or
// k_eval_cols (explicit targets in the blocks' C order), grid (groups, source chunks).
template <bool MIXED, int W>
static void column_kernel_w(tgrav_ctx* c, bool blocks, dim3 g, const tg::Tgt& T, int64_t ns,
                            int64_t chunk, const tg::Out& o) {
  if (blocks)
    tg::k_eval_grid<MIXED, W><<<g, 64 * W, 0, c->stream>>>(T, c->s4.as<double4>(), c->se2.as<double>(), ns,
                                                            chunk, c->stat.as<const double>(), o);
  else
    tg::k_eval_cols<MIXED, W><<<g, 64 * W, 0, c->stream>>>(T, c->s4.as<double4>(), c->se2.as<double>(), ns,
                                                            chunk, c->stat.as<const double>(), o);
}

static int column_kernel(tgrav_ctx* c, int mode, bool blocks, int64_t ngrp, int nsplit,
                         const tg::Tgt& T, int64_t ns, int64_t chunk, const tg::Out& o) {
  const dim3 g((unsigned)ngrp, (unsigned)nsplit);
  const bool mixed = mode == TGRAV_MIXED;
  switch (c->col_waves) {
    case 4:
      if (mixed) column_kernel_w<true, 4>(c, blocks, g, T, ns, chunk, o);
      else column_kernel_w<false, 4>(c, blocks, g, T, ns, chunk, o);
      break;
    case 2:
      if (mixed) column_kernel_w<true, 2>(c, blocks, g, T, ns, chunk, o);
      else column_kernel_w<false, 2>(c, blocks, g, T, ns, chunk, o);
      break;
    default:
      if (mixed) column_kernel_w<true, 1>(c, blocks, g, T, ns, chunk, o);
      else column_kernel_w<false, 1>(c, blocks, g, T, ns, chunk, o);
  }
  LAUNCHCHECK(blocks ? "k_eval_grid" : "k_eval_cols");
  return 0;
}

static int launch_grid(tgrav_ctx* c, int mode, const tg::Tgt& T, const tg::Out& O) {
  const int64_t nt = T.n, ns = c->ns, nblk = nt / 512;
  if (nblk > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many blocks");
  int nsplit;
  int64_t chunk;
  grid_shape(c, mode, ns, nblk, &nsplit, &chunk);
  tg::Out o = O;
  if (nsplit > 1) {
    int rc0 = c->part.ensure((size_t)nsplit * 4 * nt * sizeof(double));
    if (rc0) return rc0;
    o.part = c->part.as<double>();
    o.pstride = nt;
  } else {
    o.part = nullptr;
  }
  int rc = timed_launch(c, nt, ns, [&]() -> int {
    return column_kernel(c, mode, true, nblk, nsplit, T, ns, chunk, o);
  });
  if (rc) return rc;
  if (nsplit > 1) {
    tg::Out fin = O;
    fin.part = nullptr;
    int rc1 = reduce_chunks(c, c->part.as<double>(), nt, nsplit, nt, fin);
    if (rc1) return rc1;
  }
  return 0;
}

// Explicit targets against 8^3 source blocks: K targets per lane, tiles of 64 K targets, a
// workgroup of W = col_waves waves per tile (splitting its source blocks), source blocks split
// further over blockIdx.y.  K = 2 in the mixed kernel, 4 in the FP64 one: measured on c5's reverse
// (plan) 87.7 against 90.2 ms (mixed, K = 2 against 4) and 204.8 against 206.4 ms (FP64, K = 4
// against 2); the FP64 kernel holds each block's dz for its K targets and already needs 240 VGPRs.
#ifndef TG_GS_K_MIXED
#define TG_GS_K_MIXED 2
#endif
#ifndef TG_GS_K_FP64
#define TG_GS_K_FP64 4
#endif
template <bool MIXED>
static constexpr int gs_k() { return MIXED ? TG_GS_K_MIXED : TG_GS_K_FP64; }
static int gs_k_mode(int mode) { return mode == TGRAV_MIXED ? TG_GS_K_MIXED : TG_GS_K_FP64; }

static void gridsrc_shape(const tgrav_ctx* c, int mode, int64_t nt, int64_t nblk, int64_t* ntiles,
                          int* nsplit, int64_t* chunk) {
  const int K = gs_k_mode(mode), W = c->col_waves;
  *ntiles = (nt + K * 64 - 1) / (K * 64);
  int want = c->nsplit[mode];
  const int tuned = tuned_block_split(c, mode, "blocksrc", nt, nblk, W);
  if (tuned > 0) want = tuned;
  if (want <= 0) {
    // ~12 waves resident per CU: aim at ~16 rounds of waves, with at least 4 source blocks per
    // wave and chunk (1 when that leaves fewer waves than CUs).
    const int64_t target = 16LL * 12 * c->num_cu;
    const int64_t per = (target + *ntiles * W - 1) / (*ntiles * W);
    int64_t w = std::min<int64_t>(per, std::max<int64_t>(1, nblk / (4 * W)));
    if (*ntiles * W * w < c->num_cu) w = std::min<int64_t>(per, nblk);
    want = (int)std::min<int64_t>(w, 65535);
  }
  if (want > 65535) want = 65535;
  *chunk = (nblk + want - 1) / want;
  *nsplit = (int)((nblk + *chunk - 1) / *chunk);
}

template <bool MIXED, int W>
static void gridsrc_kernel_w(tgrav_ctx* c, dim3 g, const tg::Tgt& T, const tg::Out& o, int64_t chunk,
                             int64_t nblk) {
  tg::k_eval_gridsrc<gs_k<MIXED>(), W, MIXED><<<g, 64 * W, 0, c->stream>>>(
      T, c->sax.as<double>(), c->sbe2.as<double>(), c->sms.as<double>(), nblk, chunk,
      c->stat.as<const double>(), o);
}

static int gridsrc_kernel(tgrav_ctx* c, int mode, const tg::Tgt& T, const tg::Out& o, int64_t ntiles,
                          int nsplit, int64_t chunk, int64_t nblk) {
  // blockIdx.y rounded up to a multiple of 8: the kernel then maps the tiles of each block chunk to
  // one XCD (the padding workgroups have no chunk and return).
  const unsigned ny = nsplit >= 8 ? (unsigned)((nsplit + 7) / 8 * 8) : (unsigned)nsplit;
  if (ny > 65535) return fail(TGRAV_ELIMIT, "too many block chunks (%d)", nsplit);
  const dim3 g((unsigned)ntiles, ny);
  const bool mixed = mode == TGRAV_MIXED;
  switch (c->col_waves) {
    case 4: mixed ? gridsrc_kernel_w<true, 4>(c, g, T, o, chunk, nblk) : gridsrc_kernel_w<false, 4>(c, g, T, o, chunk, nblk); break;
    case 2: mixed ? gridsrc_kernel_w<true, 2>(c, g, T, o, chunk, nblk) : gridsrc_kernel_w<false, 2>(c, g, T, o, chunk, nblk); break;
    default: mixed ? gridsrc_kernel_w<true, 1>(c, g, T, o, chunk, nblk) : gridsrc_kernel_w<false, 1>(c, g, T, o, chunk, nblk);
  }
  LAUNCHCHECK("k_eval_gridsrc");
  return 0;
}

static int launch_gridsrc(tgrav_ctx* c, int mode, const tg::Tgt& T, const tg::Out& O) {
  const int64_t nt = T.n, nblk = c->src_nblk;
  int64_t ntiles, chunk;
  int nsplit;
  gridsrc_shape(c, mode, nt, nblk, &ntiles, &nsplit, &chunk);
  if (ntiles > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many targets");
  tg::Out o = O;
  if (nsplit > 1) {
    int rc0 = c->part.ensure((size_t)nsplit * 4 * nt * sizeof(double));
    if (rc0) return rc0;
    o.part = c->part.as<double>();
    o.pstride = nt;
  } else {
    o.part = nullptr;
  }
  int rc = timed_launch(c, nt, c->ns, [&]() -> int {
    return gridsrc_kernel(c, mode, T, o, ntiles, nsplit, chunk, nblk);
  });
  if (rc) return rc;
  if (nsplit > 1) {
    tg::Out fin = O;
    fin.part = nullptr;
    int rc1 = reduce_chunks(c, c->part.as<double>(), nt, nsplit, nt, fin);
    if (rc1) return rc1;
  }
  return 0;
}

// Largest double x with fl(sqrt(x)) < r, so that for any double r2, sqrt(r2) < r <=> r2 <= x
// (the correctly rounded sqrt is monotone); -1 if no x >= 0 qualifies (r <= 0).
static double sqrt_lt_threshold(double r) {
  if (!(r > 0.0)) return -1.0;
  double x = r * r;
  while (x > 0.0 && !(sqrt(x) < r)) x = nextafter(x, 0.0);
  while (sqrt(nextafter(x, INFINITY)) < r) x = nextafter(x, INFINITY);
  return x;
}

// FLASH kernel, isolated, explicit targets against 8^3 source blocks: K targets per lane
// (TG_FLASH_GS_K), 256 lanes, source blocks split over blockIdx.y (about 16 rounds of
// workgroups, or the chunk count of tgrav_set_config).
#ifndef TG_FLASH_GS_K
#define TG_FLASH_GS_K 2
#endif
static void flash_gridsrc_shape(const tgrav_ctx* c, int64_t nt, int64_t nblk, int64_t* ntiles,
                                int* nsplit, int64_t* chunk) {
  const int K = TG_FLASH_GS_K, NT = 256;
  *ntiles = (nt + K * NT - 1) / (K * NT);
  // As the Plummer block-source kernel: about 16 rounds of workgroups, at least 4 blocks a chunk
  // (1 when that leaves fewer workgroups than CUs).
  const int64_t target = 16LL * 3 * c->num_cu;
  const int64_t per = (target + *ntiles - 1) / *ntiles;
  int64_t want = std::min<int64_t>(per, std::max<int64_t>(1, nblk / 4));
  if (*ntiles * want < c->num_cu) want = std::min<int64_t>(per, nblk);
  if (c->nsplit[TGRAV_PLAIN_FP64] > 0) want = c->nsplit[TGRAV_PLAIN_FP64];   // tgrav_set_config
  want = std::max<int64_t>(1, std::min<int64_t>(want, 65535));
  *chunk = (nblk + want - 1) / want;
  *nsplit = (int)((nblk + *chunk - 1) / *chunk);
}

static int flash_gridsrc_kernel(tgrav_ctx* c, const tg::Tgt& T, const tg::Out& o, const tg::Flash& P,
                                int64_t ntiles, int nsplit, int64_t chunk, int64_t nblk) {
  const int K = TG_FLASH_GS_K, NT = 256;
  const dim3 g((unsigned)ntiles, (unsigned)nsplit);
  tg::k_eval_flash_gridsrc<K, NT><<<g, NT, 0, c->stream>>>(
      T, c->sax.as<double>(), c->sbe2.as<double>(), c->sms.as<double>(), nblk, chunk, P,
      c->stat.as<const double>(), o);
  LAUNCHCHECK("k_eval_flash_gridsrc");
  return 0;
}

// Periodic FLASH sums with the pairs grouped by separation (k_eval_flash_pairs), in batches of
// targets of at most kMaxPairs pairs: pair keys, a stable radix sort of them, the grouped walk
// (the timed kernel), the per-target sums.  ns <= kMaxPairs (the caller checks).
static const int64_t kMaxPairs = 1LL << 25;
static int launch_flash_pairs(tgrav_ctx* c, const tg::Tgt& T, const tg::Out& O, const tg::Flash& P) {
  const int64_t nt = T.n, ns = c->ns;
  const int64_t tb = std::max<int64_t>(1, std::min<int64_t>(nt, kMaxPairs / ns)), nmax = tb * ns;
  size_t tmp = 0;
  HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const unsigned int*)nullptr, (unsigned int*)nullptr,
                                             (const unsigned int*)nullptr, (unsigned int*)nullptr, (int)nmax, 0,
                                             3 * TG_PAIR_BITS, c->stream));
  int rc;
  if ((rc = c->pkey.ensure((size_t)nmax * 2 * sizeof(unsigned int))) ||
      (rc = c->pidx.ensure((size_t)nmax * 2 * sizeof(unsigned int))) || (rc = c->ptmp.ensure(tmp)) ||
      (rc = c->pbuf.ensure((size_t)nmax * sizeof(double4))))
    return rc;
  unsigned int *k0 = c->pkey.as<unsigned int>(), *k1 = k0 + nmax;
  unsigned int *i0 = c->pidx.as<unsigned int>(), *i1 = i0 + nmax;
  const double4* s4 = c->s4.as<double4>();
  double4* pb = c->pbuf.as<double4>();
  const double bins = (double)(1 << (TG_PAIR_BITS - 4));   // per box length
  tg::Out fin = O;
  fin.part = nullptr;
  for (int64_t t0 = 0; t0 < nt; t0 += tb) {
    const int64_t nb = std::min<int64_t>(tb, nt - t0), np = nb * ns;
    tg::k_pair_keys<<<grid1(np, 256), 256, 0, c->stream>>>(T, t0, s4, ns, np, bins / P.lx, bins / P.ly,
                                                           bins / P.lz, k0, i0);
    LAUNCHCHECK("k_pair_keys");
    size_t tmp_b = tmp;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(c->ptmp.p, tmp_b, k0, k1, i0, i1, (int)np, 0, 3 * TG_PAIR_BITS,
                                               c->stream));
    rc = timed_launch(c, nb, ns, [&]() -> int {
      const unsigned int nw = (unsigned int)((np + 63) / 64);
      if (c->nrep <= TG_PAIR_NREP) {
        if (P.quirks) tg::k_eval_flash_pairs<true, false><<<nw, 64, 0, c->stream>>>(T, t0, s4, ns, np, i1, P, pb);
        else tg::k_eval_flash_pairs<false, false><<<nw, 64, 0, c->stream>>>(T, t0, s4, ns, np, i1, P, pb);
      } else {   // nrep <= TG_PAIR_NREP_DYN (the caller checks): the table in dynamic LDS
        const size_t lds = (size_t)(2 * c->nrep + 1) * (2 * c->nrep + 1) * sizeof(tg::PairRow);
        if (P.quirks) tg::k_eval_flash_pairs<true, true><<<nw, 64, lds, c->stream>>>(T, t0, s4, ns, np, i1, P, pb);
        else tg::k_eval_flash_pairs<false, true><<<nw, 64, lds, c->stream>>>(T, t0, s4, ns, np, i1, P, pb);
      }
      LAUNCHCHECK("k_eval_flash_pairs");
      return 0;
    });
    if (rc) return rc;
    tg::k_pair_reduce<<<(unsigned int)nb, 64, 0, c->stream>>>(pb, s4, ns, t0, fin);
    LAUNCHCHECK("k_pair_reduce");
  }
  return 0;
}

// FLASH softening kernels (FP64 only).  Isolated: K = 2 targets per lane, sources split until
// there are about 16 workgroups per CU (4 waves each, so the last round of workgroups is a
// small share of the launch); periodic: K = 1 (hundreds of images per pair), about 8.
static int launch_flash(tgrav_ctx* c, int mode, const tg::Tgt& T, const tg::Out& O) {
  if (mode != TGRAV_PLAIN_FP64)
    return fail(TGRAV_EINVAL, "the FLASH softening kernels run in TGRAV_PLAIN_FP64 only");
  const int64_t nt = T.n, ns = c->ns;
  const bool per = c->periodic != 0;
  const int NT = 256, K = per ? 1 : 2;
  tg::Flash P;
  P.rsoft = c->rsoft;
  P.hinv = 2.0 / c->rsoft;
  P.h2inv = P.hinv * P.hinv;
  P.slope = 1.0 / (c->rsoft * c->rsoft * c->rsoft);
  P.soft_thr = sqrt_lt_threshold(c->rsoft);
  P.near2 = P.soft_thr * (1.0 + 1e-12);
  P.lx = c->box[0];
  P.ly = c->box[1];
  P.lz = c->box[2];
  P.ilx = per ? 1.0 / P.lx : 0.0;
  P.ily = per ? 1.0 / P.ly : 0.0;
  P.ilz = per ? 1.0 / P.lz : 0.0;
  P.cut_thr = per ? sqrt_lt_threshold((double)c->nrep * std::min(P.lx, std::min(P.ly, P.lz))) : -1.0;
  P.type = c->soft;
  P.quirks = c->soft_quirks;
  P.nrep = per ? c->nrep : 0;
  // Pairs reaching the refined 1/sqrt have r2 >= rsoft^2 (the rest take the softened
  // branch); the FP32 seed is valid from 1e-30 up, and up to the tile's r^2 bound (in the kernels).
  P.s32 = TG_FP64_SEED32 && c->rsoft * c->rsoft > TG_F32_TINY_E2;
  P.reach = per ? (double)(c->nrep + 1) * std::max(P.lx, std::max(P.ly, P.lz)) : 0.0;
  // Periodic sums with the pairs grouped by separation, where the row table fits in LDS (nrep <= 7
  // static, nrep <= 10 sized at launch), in batches of targets of at most 2^25 pairs (1.3 GB of pair
  // arrays); tgrav_set_periodic(on = 2) keeps the per-target walk.
  if (per && c->periodic == 1 && c->nrep <= TG_PAIR_NREP_DYN && nt > 0 && ns > 0 && ns <= kMaxPairs &&
      nt <= 0x7fffffffLL)
    return launch_flash_pairs(c, T, O, P);
  // Sources given as 8^3 blocks take the block-source kernel.  Explicit sources that may be
  // C-ordered 8^3 blocks (flagged on the device when they were set) launch it too, gated against
  // the generic kernel on that flag, as in the Plummer path.
  const bool isolated_blocks = !per && !T.blocks && c->block_kernel;
  const bool srcb = isolated_blocks && c->src_nblk > 0;
  const bool cand = isolated_blocks && !srcb && c->src_cols > 0;
  int64_t ntiles_s = 0, chunk_s = 0;
  int nsplit_s = 1;
  const int64_t nblk_s = srcb ? c->src_nblk : c->src_cols;
  if (srcb || cand) flash_gridsrc_shape(c, nt, nblk_s, &ntiles_s, &nsplit_s, &chunk_s);
  if (ntiles_s > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many targets (%lld)", (long long)nt);
  const int64_t ntiles = (nt + (int64_t)K * NT - 1) / ((int64_t)K * NT);
  if (ntiles > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many targets (%lld)", (long long)nt);
  const int64_t target_wg = (per ? 8LL : 16LL) * c->num_cu;
  int64_t want = 1;
  if (ntiles < target_wg / 2) {
    want = (target_wg + ntiles - 1) / ntiles;
    want = std::min<int64_t>(want, per ? ns : (ns + 63) / 64);
    want = std::min<int64_t>(want, (int64_t)(((size_t)1 << 31) / ((size_t)32 * nt)));
    want = std::max<int64_t>(1, std::min<int64_t>(want, 65535));
  }
  const int64_t chunk = (ns + want - 1) / want;
  const int nsplit = srcb ? 1 : (int)((ns + chunk - 1) / chunk);
  const int nsp = std::max(nsplit, (srcb || cand) ? nsplit_s : 1);
  if (nsp > 1) {
    int rc0 = c->part.ensure((size_t)nsp * 4 * nt * sizeof(double));
    if (rc0) return rc0;
  }
  tg::Out o = O, os = O;
  o.part = nsplit > 1 ? c->part.as<double>() : nullptr;
  os.part = nsplit_s > 1 ? c->part.as<double>() : nullptr;
  o.pstride = os.pstride = nt;
  if (cand) {
    int rc0 = c->colflag.ensure(sizeof(int));
    if (rc0) return rc0;
    int* flag = c->colflag.as<int>();
    tg::k_eval_gate<<<1, 64, 0, c->stream>>>(flag, c->scolflag.as<const int>(), 0);
    LAUNCHCHECK("k_eval_gate");
    o.gate = os.gate = flag;
    o.gate_on = 0;
    os.gate_on = 2;
  }
  int rc = timed_launch(c, nt, ns, [&]() -> int {
    if (srcb || cand) {
      int r = flash_gridsrc_kernel(c, T, os, P, ntiles_s, nsplit_s, chunk_s, nblk_s);
      if (r) return r;
    }
    if (srcb) return 0;
    const dim3 g((unsigned)ntiles, (unsigned)nsplit);
    const double4* s4 = c->s4.as<double4>();
    const double* sst = c->stat.as<const double>();
    if (per) {
      if (P.quirks)
        tg::k_eval_flash<1, 256, true, true><<<g, NT, 0, c->stream>>>(T, s4, ns, chunk, P, sst, o);
      else
        tg::k_eval_flash<1, 256, true, false><<<g, NT, 0, c->stream>>>(T, s4, ns, chunk, P, sst, o);
    } else {
      tg::k_eval_flash<2, 256, false><<<g, NT, 0, c->stream>>>(T, s4, ns, chunk, P, sst, o);
    }
    LAUNCHCHECK("k_eval_flash");
    return 0;
  });
  if (rc) return rc;
  if ((srcb || cand) && nsplit_s > 1) {
    tg::Out fin = os;
    fin.part = nullptr;
    int rc1 = reduce_chunks(c, c->part.as<double>(), nt, nsplit_s, nt, fin);
    if (rc1) return rc1;
  }
  if (!srcb && nsplit > 1) {
    tg::Out fin = o;
    fin.part = nullptr;
    int rc1 = reduce_chunks(c, c->part.as<double>(), nt, nsplit, nt, fin);
    if (rc1) return rc1;
  }
  return 0;
}

// Particle side of the fused passes: sums the ntiles cell-tile partials per particle in a
// fixed order.  Above 64 tiles the sum runs in two levels (64 groups of consecutive tiles, then
// the groups in order), so the partial buffer is read by ~64x more workgroups than particles
// alone would give.
static int reduce_particle_partials(tgrav_ctx* c, int64_t ntiles, int64_t np, const tg::Out& PO) {
  const double* src = c->bppart.as<double>();
  int64_t nt = ntiles;
  if (ntiles > 64) {
    const int G = 64;
    const int per = (int)((ntiles + G - 1) / G);
    const int groups = (int)((ntiles + per - 1) / per);
    int rc = c->bpgrp.ensure((size_t)groups * 4 * np * sizeof(double));
    if (rc) return rc;
    tg::k_bidir_group<<<dim3(grid1(np, 256), (unsigned)groups), 256, 0, c->stream>>>(
        src, (int)ntiles, per, np, c->bpgrp.as<double>(), PO);
    LAUNCHCHECK("k_bidir_group");
    src = c->bpgrp.as<double>();
    nt = groups;
  }
  tg::k_bidir_reduce<<<grid1(np, 256), 256, 0, c->stream>>>(src, (int)nt, np, PO);
  LAUNCHCHECK("k_bidir_reduce");
  return 0;
}

// Core launch: targets T (device), outputs O (device), on c->stream.
// Variant of the generic kernels for ns sources onto nt targets: the tune cache's entry for this
// size bucket (with its source split in *force_split), else the context's choice.
static const Variant* generic_variant(const tgrav_ctx* c, int mode, int64_t ns, int64_t nt, int* force_split) {
  const Variant* v = variant_of(mode, c->variant[mode]);
  *force_split = -1;
  if (!c->user_cfg[mode] && !c->tune.empty()) {
    char key[256];
    snprintf(key, sizeof key, "%s|eval|%s|s%d|t%d", c->devkey.c_str(), mode_name(mode),
             size_bucket((double)ns), size_bucket((double)nt));
    const TuneEntry* t = tune_lookup(c, key);
    if (t && t->variant >= 0 && variant_of(mode, t->variant)) {
      v = variant_of(mode, t->variant);
      *force_split = t->nsplit;
    }
  }
  return v;
}

static int launch_eval(tgrav_ctx* c, int mode, const tg::Tgt& T, const tg::Out& O) {
  const int64_t nt = T.n, ns = c->ns;
  if (nt == 0) return 0;
  // Every pointer a kernel dereferences must be set before anything is launched.
  if (T.blocks ? (!T.ox || !T.oy || !T.oz || !T.bdx) : (!T.x || !T.y || !T.z))
    return fail(TGRAV_EINVAL, "internal: target arrays not bound");
  if (O.vx && (!O.vy || !O.vz)) return fail(TGRAV_EINVAL, "internal: partial velocity set");
  if (ns > 0 && (!c->s4.p || !c->se2.p || !c->stat.p))
    return fail(TGRAV_EINVAL, "internal: source buffers missing");
  if (c->soft == TGRAV_SOFT_PLUMMER && c->periodic)
    return fail(TGRAV_EINVAL, "the periodic image sum needs a FLASH softening kernel (tgrav_set_softening)");
  if (c->soft != TGRAV_SOFT_PLUMMER && mode != TGRAV_PLAIN_FP64)
    return fail(TGRAV_EINVAL, "the FLASH softening kernels run in TGRAV_PLAIN_FP64 only");
  if (ns == 0) {   // no sources: zero field, kick adds nothing
    tg::k_fill<<<grid1(nt, 256), 256, 0, c->stream>>>(nt, 0.0, O.ax, O.ay, O.az, O.phi);
    LAUNCHCHECK("k_fill");
    return 0;
  }
  if (c->soft != TGRAV_SOFT_PLUMMER) return launch_flash(c, mode, T, O);
  if (T.blocks && T.nb == 8 && c->block_kernel) return launch_grid(c, mode, T, O);
  if (!T.blocks && c->src_nblk > 0 && c->block_kernel) return launch_gridsrc(c, mode, T, O);
  int force_split = -1;
  const Variant* v = generic_variant(c, mode, ns, nt, &force_split);
  int64_t ntiles, chunk;
  int nsplit;
  choose_split(c, mode, v, nt, ns, &ntiles, &nsplit, &chunk, force_split);
  if (ntiles > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many targets (%lld)", (long long)nt);
  // Explicit targets in the C order of 8^3 leaf-block meshgrids (the reference's call shape,
  // .deprecated/bridge_multiples.py:566-575) run the column kernel.  Whether they have that
  // structure is decided on the device (k_check_cols sets a flag); both kernels are launched
  // and the flag lets exactly one of them (and its chunk reduction) run, so nothing waits for
  // the host.
  // Likewise explicit sources that are C-ordered 8^3 blocks (k_detect_source_blocks flagged
  // them when they were set) run the block-source kernel: gate value 2 wins over the column
  // test (1) and the generic kernel (0).
  const bool cols = c->block_kernel && !T.blocks && nt >= 512;
  const bool srcb = c->block_kernel && !T.blocks && c->src_cols > 0;
  const int64_t ngrp = (nt + 511) / 512;
  int nsplit_c = 1, nsplit_s = 1;
  int64_t chunk_c = 0, chunk_s = 0, ntiles_s = 0;
  if (cols) {
    if (ngrp > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many targets (%lld)", (long long)nt);
    grid_shape(c, mode, ns, ngrp, &nsplit_c, &chunk_c);
  }
  if (srcb) gridsrc_shape(c, mode, nt, c->src_cols, &ntiles_s, &nsplit_s, &chunk_s);
  const int nsp = std::max(nsplit, std::max(cols ? nsplit_c : 1, srcb ? nsplit_s : 1));
  if (nsp > 1) {
    int rc0 = c->part.ensure((size_t)nsp * 4 * nt * sizeof(double));
    if (rc0) return rc0;
  }
  tg::Out o = O, oc = O, os = O;
  o.part = nsplit > 1 ? c->part.as<double>() : nullptr;
  oc.part = nsplit_c > 1 ? c->part.as<double>() : nullptr;
  os.part = nsplit_s > 1 ? c->part.as<double>() : nullptr;
  o.pstride = oc.pstride = os.pstride = nt;
  if (cols || srcb) {
    int rc0 = c->colflag.ensure(sizeof(int));
    if (rc0) return rc0;
    int* flag = c->colflag.as<int>();
    tg::k_eval_gate<<<1, 64, 0, c->stream>>>(flag, srcb ? c->scolflag.as<const int>() : nullptr,
                                              cols ? 1 : 0);
    LAUNCHCHECK("k_eval_gate");
    if (cols) {
      tg::k_check_cols<<<grid1(nt, 256), 256, 0, c->stream>>>(T, flag);
      LAUNCHCHECK("k_check_cols");
    }
    o.gate = oc.gate = os.gate = flag;
    o.gate_on = 0;
    oc.gate_on = 1;
    os.gate_on = 2;
  }
  int rc = timed_launch(c, nt, ns, [&]() -> int {
    if (srcb) {
      int r = gridsrc_kernel(c, mode, T, os, ntiles_s, nsplit_s, chunk_s, c->src_cols);
      if (r) return r;
    }
    if (cols) {
      int r = column_kernel(c, mode, false, ngrp, nsplit_c, T, ns, chunk_c, oc);
      if (r) return r;
    }
    v->fn(dim3((unsigned)ntiles, (unsigned)nsplit), dim3(v->nt), c->stream, T, c->s4.as<double4>(),
          c->se2.as<double>(), ns, chunk, c->stat.as<const double>(), o);
    LAUNCHCHECK(mode == TGRAV_MIXED ? "k_eval_mixed" : "k_eval_fp64");
    return 0;
  });
  if (rc) return rc;
  if (srcb && nsplit_s > 1) {
    tg::Out fin = os;
    fin.part = nullptr;
    int rc1 = reduce_chunks(c, c->part.as<double>(), nt, nsplit_s, nt, fin);
    if (rc1) return rc1;
  }
  if (cols && nsplit_c > 1) {
    tg::Out fin = oc;
    fin.part = nullptr;
    int rc1 = reduce_chunks(c, c->part.as<double>(), nt, nsplit_c, nt, fin);
    if (rc1) return rc1;
  }
  if (nsplit > 1) {
    tg::Out fin = o;
    fin.part = nullptr;
    int rc1 = reduce_chunks(c, c->part.as<double>(), nt, nsplit, nt, fin);
    if (rc1) return rc1;
  }
  return 0;
}

// ------------------------------------------------------------------- C ABI -------
extern "C" {

int tgrav_version(void) { return TGRAV_VERSION; }

const char* tgrav_last_error(void) { return g_err; }

int tgrav_tune_cache_entries(const char* path) {
  if (!path) return fail(TGRAV_EINVAL, "path is NULL");
  std::string text;
  if (read_file(path, &text)) return fail(TGRAV_EINVAL, "cannot read %s", path);
  const int n = parse_tune_cache(text, nullptr);
  if (n < 0) return fail(TGRAV_EINVAL, "%s is not a tune cache", path);
  return n;
}

int tgrav_create(tgrav_ctx** out, int device, const char* tune_cache_path) {
  if (!out) return fail(TGRAV_EINVAL, "tgrav_create: out is NULL");
  *out = nullptr;
  std::map<std::string, TuneEntry> tune;
  if (tune_cache_path) {
    std::string text;
    if (read_file(tune_cache_path, &text))
      return fail(TGRAV_EINVAL, "tgrav_create: cannot read tune cache %s", tune_cache_path);
    if (parse_tune_cache(text, &tune) < 0)
      return fail(TGRAV_EINVAL, "tgrav_create: %s is not a tune cache", tune_cache_path);
  }
  int ndev = 0;
  HIPCHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(TGRAV_EINVAL, "tgrav_create: device %d out of range (%d devices)", device, ndev);
  tgrav_ctx* c = new tgrav_ctx();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) {
      c->num_cu = prop.multiProcessorCount;
      std::string arch(prop.gcnArchName);
      arch = arch.substr(0, arch.find(':'));
      c->devkey = arch + "/" + std::to_string(c->num_cu) + "cu";
    }
    e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  }
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&c->host_flag, kHostWords * sizeof(unsigned long long), hipHostMallocDefault);
  if (e != hipSuccess) {
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return fail(TGRAV_EHIP, "tgrav_create: %s", hipGetErrorString(e));
  }
  const int rs = scratch_init(device);
  if (rs) {
    (void)hipHostFree(c->host_flag);
    (void)hipStreamDestroy(c->own);
    delete c;
    return rs;
  }
  c->stream = c->own;
  c->tune.swap(tune);
  *out = c;
  return 0;
}

int tgrav_destroy(tgrav_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->s4.release();
  c->se2.release();
  for (auto& b : c->in) b.release();
  for (auto& b : c->outb) b.release();
  for (auto& b : c->vel) b.release();
  c->part.release();
  c->stat.release();
  c->rstat.release();
  c->flag.release();
  c->colflag.release();
  c->sbe2.release();
  c->sms.release();
  c->sax.release();
  c->scolflag.release();
  c->bp4.release();
  c->bpe2.release();
  c->bstat.release();
  c->bppart.release();
  c->bpgrp.release();
  for (auto& b : c->pm) b.release();
  for (auto& b : c->pm_acc) b.release();
  c->pm_stat.release();
  c->pkey.release();
  c->pidx.release();
  c->ptmp.release();
  c->pbuf.release();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->host_flag) (void)hipHostFree(c->host_flag);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
  return 0;
}

int tgrav_set_stream(tgrav_ctx* c, void* s) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  c->stream = (hipStream_t)s;
  return 0;
}

int tgrav_use_own_stream(tgrav_ctx* c) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  c->stream = c->own;
  return 0;
}

void* tgrav_get_stream(tgrav_ctx* c) { return c ? (void*)c->stream : nullptr; }

int tgrav_synchronize(tgrav_ctx* c) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc = set_device(c);
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int tgrav_set_constant(tgrav_ctx* c, double G) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  if (!isfinite(G)) return fail(TGRAV_EINVAL, "G must be finite");
  c->G = G;
  return 0;
}

int tgrav_set_validation(tgrav_ctx* c, int on) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  c->validate = on ? 1 : 0;
  return 0;
}

int tgrav_num_variants(int mode) {
  if (mode == TGRAV_PLAIN_FP64) return kNumFp64;
  if (mode == TGRAV_MIXED) return kNumMixed;
  return fail(TGRAV_EINVAL, "bad mode %d", mode);
}

int tgrav_variant_info(int mode, int variant, int* k, int* threads, int* ts) {
  const Variant* v = variant_of(mode, variant);
  if (!v) return fail(TGRAV_EINVAL, "no variant %d for mode %d", variant, mode);
  if (k) *k = v->k;
  if (threads) *threads = v->nt;
  if (ts) *ts = v->ts;
  return 0;
}

int tgrav_set_config(tgrav_ctx* c, int mode, int variant, int nsplit) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc = check_mode(mode);
  if (rc) return rc;
  if (variant >= 0) {
    if (!variant_of(mode, variant)) return fail(TGRAV_EINVAL, "no variant %d for mode %d", variant, mode);
    c->variant[mode] = variant;
  }
  if (nsplit >= 0) c->nsplit[mode] = nsplit;
  c->user_cfg[mode] = true;
  return 0;
}

int tgrav_set_sources(tgrav_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                      const double* m, const double* eps, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc = check_kind(kind);
  if (rc) return rc;
  if (n < 0) return fail(TGRAV_EINVAL, "number of sources must be >= 0 (got %lld)", (long long)n);
  if (n > 0 && (!x || !y || !z || !m))
    return fail(TGRAV_EINVAL, "source x, y, z and m are required");
  if ((rc = set_device(c))) return rc;
  c->ns = -1;
  c->src_nblk = 0;
  c->src_cols = 0;
  if (n == 0) {
    c->ns = 0;
    return 0;
  }
  if ((rc = c->s4.ensure((size_t)n * sizeof(double4)))) return rc;
  if ((rc = c->se2.ensure((size_t)n * sizeof(double)))) return rc;
  if ((rc = c->stat.ensure(tg::ST_WORDS * sizeof(unsigned long long)))) return rc;
  const double* a[5] = {x, y, z, m, eps};
  if (kind == TGRAV_HOST) {
    if (c->validate)
      for (int q = 0; q < 5; ++q)
        if (!host_finite(n, a[q])) return fail(TGRAV_ENONFINITE, "non-finite value in source array %d", q);
    if ((rc = stage_in(c, 0, n, a, 5))) return rc;
  }
  tg::k_reset_stats<<<1, 64, 0, c->stream>>>(c->stat.as<unsigned long long>());
  LAUNCHCHECK("k_reset_stats");
  tg::k_pack_sources<<<grid1(n, 256), 256, 0, c->stream>>>(
      n, a[0], a[1], a[2], a[3], a[4], c->s4.as<double4>(), c->se2.as<double>(),
      c->stat.as<unsigned long long>(), kind == TGRAV_DEVICE && c->validate);
  LAUNCHCHECK("k_pack_sources");
  const int64_t nb8 = n / 512;
  if (n % 512 == 0) {   // C-ordered 8^3 blocks? decided on the device (k_detect_source_blocks)
    if ((rc = c->sax.ensure((size_t)nb8 * 24 * sizeof(double))) ||
        (rc = c->sbe2.ensure((size_t)nb8 * sizeof(double))) || (rc = c->sms.ensure((size_t)n * 8)) ||
        (rc = c->scolflag.ensure(sizeof(int))))
      return rc;
    tg::k_set_flag<<<1, 64, 0, c->stream>>>(c->scolflag.as<int>(), 1);
    LAUNCHCHECK("k_set_flag");
    tg::k_detect_source_blocks<<<grid1(n, 256), 256, 0, c->stream>>>(
        nb8, c->s4.as<double4>(), c->se2.as<double>(), c->sax.as<double>(), c->sbe2.as<double>(),
        c->sms.as<double>(), c->scolflag.as<int>());
    LAUNCHCHECK("k_detect_source_blocks");
  }
  if (kind == TGRAV_DEVICE) {
    if (c->validate) {
      bool bad = false;
      if ((rc = read_flag(c, c->stat.as<unsigned long long>() + tg::ST_BAD, &bad))) return rc;
      if (bad) return fail(TGRAV_ENONFINITE, "non-finite value in a source array");
    }
  } else {
    HIPCHECK(hipStreamSynchronize(c->stream));   // host arrays may be freed on return
  }
  c->ns = n;
  c->src_cols = n % 512 == 0 ? nb8 : 0;
  return 0;
}

// Points the target descriptor at the (device) arrays tin[].
static void bind_targets(tg::Tgt& T, const double* const* tin, int ntin) {
  if (T.blocks) {
    T.ox = tin[0];
    T.oy = tin[1];
    T.oz = tin[2];
    T.bdx = tin[3];
    T.beps = ntin > 4 ? tin[4] : nullptr;
  } else {
    T.x = tin[0];
    T.y = tin[1];
    T.z = tin[2];
    T.eps = tin[3];
  }
}

int tgrav_set_source_blocks(tgrav_ctx* c, int64_t nblk, int nb, const double* ox,
                            const double* oy, const double* oz, const double* dx, const double* m,
                            const double* eps, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc = check_kind(kind);
  if (rc) return rc;
  if (nblk < 0) return fail(TGRAV_EINVAL, "number of blocks must be >= 0");
  if (nb != 8) return fail(TGRAV_EINVAL, "source blocks must have 8^3 cells (got nb = %d)", nb);
  if (nblk > 0 && (!ox || !oy || !oz || !dx || !m))
    return fail(TGRAV_EINVAL, "block origins, widths and cell masses are required");
  if ((rc = set_device(c))) return rc;
  c->ns = -1;
  c->src_nblk = 0;
  c->src_cols = 0;
  if (nblk == 0) {
    c->ns = 0;
    return 0;
  }
  const int64_t n = nblk * 512;
  const double* a[6] = {ox, oy, oz, dx, eps, m};
  if (kind == TGRAV_HOST) {
    if (c->validate) {
      for (int q = 0; q < 5; ++q)
        if (!host_finite(nblk, a[q])) return fail(TGRAV_ENONFINITE, "non-finite block value");
      if (!host_finite(n, m)) return fail(TGRAV_ENONFINITE, "non-finite cell mass");
    }
    if ((rc = stage_in(c, 0, nblk, a, 5)) || (rc = stage_in(c, 5, n, a + 5, 1))) return rc;
  } else if (c->validate) {
    bool bad = false;
    if ((rc = device_check_finite(c, nblk, a, 5, &bad))) return rc;
    if (!bad && (rc = device_check_finite(c, n, a + 5, 1, &bad))) return rc;
    if (bad) return fail(TGRAV_ENONFINITE, "non-finite value in a source block array");
  }
  if ((rc = c->s4.ensure((size_t)n * sizeof(double4))) || (rc = c->se2.ensure((size_t)n * 8)) ||
      (rc = c->stat.ensure(tg::ST_WORDS * sizeof(unsigned long long))) ||
      (rc = c->sbe2.ensure((size_t)nblk * 8)) || (rc = c->sms.ensure((size_t)n * 8)) ||
      (rc = c->sax.ensure((size_t)nblk * 24 * sizeof(double))))
    return rc;
  tg::k_reset_stats<<<1, 64, 0, c->stream>>>(c->stat.as<unsigned long long>());
  LAUNCHCHECK("k_reset_stats");
  tg::k_pack_source_blocks<<<grid1(n, 256), 256, 0, c->stream>>>(
      nblk, a[0], a[1], a[2], a[3], a[5], a[4], c->s4.as<double4>(), c->se2.as<double>(),
      c->sbe2.as<double>(), c->sax.as<double>(), c->sms.as<double>(), c->stat.as<unsigned long long>());
  LAUNCHCHECK("k_pack_source_blocks");
  if (kind == TGRAV_HOST) HIPCHECK(hipStreamSynchronize(c->stream));
  c->ns = n;
  c->src_nblk = nblk;
  return 0;
}

// Shared body of tgrav_eval / tgrav_eval_kick / tgrav_eval_blocks.
static int eval_common(tgrav_ctx* c, tg::Tgt T, const double* tin[5], int ntin, double* outs[4],
                       double dt, double* vel[3], int mode, int kind) {
  int rc;
  const int64_t n = T.n;
  const int64_t nin = T.blocks ? n / ((int64_t)T.nb * T.nb * T.nb) : n;
  const bool kick = vel[0] != nullptr;
  if (kind == TGRAV_HOST) {
    if (c->validate)
      for (int q = 0; q < ntin; ++q)
        if (!host_finite(nin, tin[q]))
          return fail(TGRAV_ENONFINITE, "non-finite value in target array %d", q);
    if ((rc = stage_in(c, 0, nin, tin, ntin))) return rc;
    bind_targets(T, tin, ntin);
    if (mode == TGRAV_MIXED && c->validate && (rc = eval_range_check(c, T, nin))) return rc;
    double* dout[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int q = 0; q < 4; ++q)
      if (outs[q]) {
        if ((rc = c->outb[q].ensure((size_t)n * 8))) return rc;
        dout[q] = c->outb[q].as<double>();
      }
    double* dvel[3] = {nullptr, nullptr, nullptr};
    if (kick)
      for (int q = 0; q < 3; ++q) {
        if ((rc = c->vel[q].ensure((size_t)n * 8))) return rc;
        dvel[q] = c->vel[q].as<double>();
        HIPCHECK(hipMemcpyAsync(dvel[q], vel[q], (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
      }
    tg::Out O{dout[0], dout[1], dout[2], dout[3], dvel[0], dvel[1], dvel[2], dt, nullptr, 0, c->G,
              outs[3] == nullptr};
    if ((rc = launch_eval(c, mode, T, O))) return rc;
    for (int q = 0; q < 4; ++q)
      if (outs[q])
        HIPCHECK(hipMemcpyAsync(outs[q], dout[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    if (kick)
      for (int q = 0; q < 3; ++q)
        HIPCHECK(hipMemcpyAsync(vel[q], dvel[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return 0;
  }
  if (c->validate && ntin > 0) {
    bool bad = false;
    if ((rc = device_check_finite(c, nin, tin, ntin, &bad))) return rc;
    if (bad) return fail(TGRAV_ENONFINITE, "non-finite value in a target array");
  }
  bind_targets(T, tin, ntin);
  if (mode == TGRAV_MIXED && c->validate && (rc = eval_range_check(c, T, nin))) return rc;
  tg::Out O{outs[0], outs[1], outs[2], outs[3], vel[0], vel[1], vel[2], dt, nullptr, 0, c->G,
            outs[3] == nullptr};
  return launch_eval(c, mode, T, O);
}

static int eval_prologue(tgrav_ctx* c, int64_t n, int mode, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc;
  if ((rc = check_mode(mode)) || (rc = check_kind(kind))) return rc;
  if (n < 0) return fail(TGRAV_EINVAL, "number of points must be >= 0 (got %lld)", (long long)n);
  if (c->ns < 0) return fail(TGRAV_ENOSOURCES, "no source set: call tgrav_set_sources first");
  return set_device(c);
}

int tgrav_eval(tgrav_ctx* c, int64_t n, const double* eps_t, const double* x, const double* y,
               const double* z, double* ax, double* ay, double* az, double* phi, int mode, int kind) {
  int rc = eval_prologue(c, n, mode, kind);
  if (rc) return rc;
  if (n == 0) return 0;
  if (!x || !y || !z) return fail(TGRAV_EINVAL, "target x, y and z are required");
  tg::Tgt T{};
  T.n = n;
  const double* tin[5] = {x, y, z, eps_t, nullptr};
  double* outs[4] = {ax, ay, az, phi};
  double* vel[3] = {nullptr, nullptr, nullptr};
  return eval_common(c, T, tin, 4, outs, 0.0, vel, mode, kind);
}

int tgrav_eval_kick(tgrav_ctx* c, int64_t n, const double* eps_t, const double* x,
                    const double* y, const double* z, double dt, double* vx, double* vy,
                    double* vz, double* ax, double* ay, double* az, double* phi, int mode,
                    int kind) {
  int rc = eval_prologue(c, n, mode, kind);
  if (rc) return rc;
  if (n == 0) return 0;
  if (!x || !y || !z) return fail(TGRAV_EINVAL, "target x, y and z are required");
  if (!vx || !vy || !vz) return fail(TGRAV_EINVAL, "velocities vx, vy and vz are required");
  if (!isfinite(dt)) return fail(TGRAV_EINVAL, "dt must be finite");
  tg::Tgt T{};
  T.n = n;
  const double* tin[5] = {x, y, z, eps_t, nullptr};
  double* outs[4] = {ax, ay, az, phi};
  double* vel[3] = {vx, vy, vz};
  return eval_common(c, T, tin, 4, outs, dt, vel, mode, kind);
}

int tgrav_eval_blocks(tgrav_ctx* c, int64_t nblk, int nb, const double* ox, const double* oy,
                      const double* oz, const double* dx, const double* eps_b, double* ax,
                      double* ay, double* az, double* phi, double dt, double* vx, double* vy,
                      double* vz, int mode, int kind) {
  int rc = eval_prologue(c, nblk, mode, kind);
  if (rc) return rc;
  if (nb <= 0 || nb > 64) return fail(TGRAV_EINVAL, "cells per block edge must be in 1..64 (got %d)", nb);
  if (nblk == 0) return 0;
  if (!ox || !oy || !oz || !dx) return fail(TGRAV_EINVAL, "block origins and cell widths are required");
  const bool kick = vx || vy || vz;
  if (kick && (!vx || !vy || !vz)) return fail(TGRAV_EINVAL, "give all of vx, vy, vz or none");
  if (kick && !isfinite(dt)) return fail(TGRAV_EINVAL, "dt must be finite");
  tg::Tgt T{};
  T.blocks = 1;
  T.nb = nb;
  T.n = nblk * (int64_t)nb * nb * nb;
  const double* tin[5] = {ox, oy, oz, dx, eps_b};
  double* outs[4] = {ax, ay, az, phi};
  double* vel[3] = {kick ? vx : nullptr, kick ? vy : nullptr, kick ? vz : nullptr};
  return eval_common(c, T, tin, 5, outs, dt, vel, mode, kind);
}

// The allocation-free variant of tgrav_eval (device pointers, sources passed in): packed sources,
// their statistics and the chunk partial sums live in the fixed per-device array `scratch`
// (layout: ST_WORDS statistics words padded to 256 B, then {x,y,z,m} and eps^2 of the sources,
// then the partials); the kernels and the launch shape are tgrav_eval's generic ones.
int tgrav_eval_static(tgrav_ctx* c, int64_t ns, const double* sx, const double* sy, const double* sz,
                      const double* sm, const double* se, int64_t nt, const double* eps_t,
                      const double* x, const double* y, const double* z, double* ax, double* ay,
                      double* az, double* phi, int mode) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc = check_mode(mode);
  if (rc) return rc;
  if (ns < 0 || nt < 0) return fail(TGRAV_EINVAL, "sizes must be >= 0");
  if (ns > 0 && (!sx || !sy || !sz || !sm)) return fail(TGRAV_EINVAL, "source x, y, z and m are required");
  if (nt > 0 && (!x || !y || !z)) return fail(TGRAV_EINVAL, "target x, y and z are required");
  if (c->soft != TGRAV_SOFT_PLUMMER || c->periodic)
    return fail(TGRAV_EINVAL, "tgrav_eval_static uses the Plummer kernel (tgrav_set_softening(ctx, 0, 0, 0))");
  if (c->device < 0 || c->device >= TG_MAX_DEVICES || !g_scratch_ptr[c->device])
    return fail(TGRAV_EINVAL, "no scratch array for device %d", c->device);
  const size_t head = 256, src_bytes = ((size_t)ns * (sizeof(double4) + sizeof(double)) + 255) / 256 * 256;
  if ((double)head + (double)src_bytes > (double)TG_STATIC_SCRATCH_BYTES)
    return fail(TGRAV_ELIMIT, "tgrav_eval_static: %lld sources need %.0f bytes of scratch, more than its %llu",
                (long long)ns, (double)head + (double)src_bytes, TG_STATIC_SCRATCH_BYTES);
  if ((rc = set_device(c))) return rc;
  if (nt == 0) return 0;
  if (ns == 0) {   // no sources: zero field
    tg::k_fill<<<grid1(nt, 256), 256, 0, c->stream>>>(nt, 0.0, ax, ay, az, phi);
    LAUNCHCHECK("k_fill");
    return 0;
  }
  unsigned char* base = (unsigned char*)g_scratch_ptr[c->device];
  unsigned long long* stat = (unsigned long long*)base;
  double4* s4 = (double4*)(base + head);
  double* se2 = (double*)(base + head + (size_t)ns * sizeof(double4));
  double* part = (double*)(base + head + src_bytes);
  const size_t part_cap = TG_STATIC_SCRATCH_BYTES - head - src_bytes;
  int force_split = -1;
  const Variant* v = generic_variant(c, mode, ns, nt, &force_split);
  int64_t ntiles, chunk;
  int nsplit;
  choose_split(c, mode, v, nt, ns, &ntiles, &nsplit, &chunk, force_split);
  if (ntiles > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many targets (%lld)", (long long)nt);
  const double per_split = 4.0 * sizeof(double) * (double)nt;
  if (nsplit > 1 && nsplit * per_split > (double)part_cap) {   // fewer chunks, so the partials fit
    const int64_t most = (int64_t)((double)part_cap / per_split);
    choose_split(c, mode, v, nt, ns, &ntiles, &nsplit, &chunk, (int)std::max<int64_t>(1, std::min<int64_t>(most, 65535)));
    if (nsplit > 1 && nsplit * per_split > (double)part_cap)
      choose_split(c, mode, v, nt, ns, &ntiles, &nsplit, &chunk, 1);
  }
  tg::Tgt T{};
  T.x = x; T.y = y; T.z = z; T.eps = eps_t; T.n = nt;
  tg::Out O{ax, ay, az, phi, nullptr, nullptr, nullptr, 0.0, nullptr, 0, c->G, phi == nullptr};
  tg::Out o = O;
  o.part = nsplit > 1 ? part : nullptr;
  o.pstride = nt;
  tg::k_reset_stats<<<1, 64, 0, c->stream>>>(stat);
  LAUNCHCHECK("k_reset_stats");
  tg::k_pack_sources<<<grid1(ns, 256), 256, 0, c->stream>>>(ns, sx, sy, sz, sm, se, s4, se2, stat, 0);
  LAUNCHCHECK("k_pack_sources");
  v->fn(dim3((unsigned)ntiles, (unsigned)nsplit), dim3(v->nt), c->stream, T, s4, se2, ns, chunk,
        (const double*)stat, o);
  LAUNCHCHECK(mode == TGRAV_MIXED ? "k_eval_mixed" : "k_eval_fp64");
  if (nsplit > 1) return reduce_chunks(c, part, nt, nsplit, nt, O);
  return 0;
}

int tgrav_kick(tgrav_ctx* c, int64_t n, double dt, const double* ax, const double* ay,
               const double* az, double* vx, double* vy, double* vz, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc = check_kind(kind);
  if (rc) return rc;
  if (n < 0) return fail(TGRAV_EINVAL, "n must be >= 0");
  if (n == 0) return 0;
  if (!ax || !ay || !az || !vx || !vy || !vz) return fail(TGRAV_EINVAL, "all six arrays are required");
  if (!isfinite(dt)) return fail(TGRAV_EINVAL, "dt must be finite");
  if ((rc = set_device(c))) return rc;
  if (kind == TGRAV_DEVICE) {
    tg::k_kick<<<grid1(n, 256), 256, 0, c->stream>>>(n, dt, ax, ay, az, vx, vy, vz);
    LAUNCHCHECK("k_kick");
    return 0;
  }
  const double* a[3] = {ax, ay, az};
  if ((rc = stage_in(c, 0, n, a, 3))) return rc;
  double* dv[3];
  double* hv[3] = {vx, vy, vz};
  for (int q = 0; q < 3; ++q) {
    if ((rc = c->vel[q].ensure((size_t)n * 8))) return rc;
    dv[q] = c->vel[q].as<double>();
    HIPCHECK(hipMemcpyAsync(dv[q], hv[q], (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
  }
  tg::k_kick<<<grid1(n, 256), 256, 0, c->stream>>>(n, dt, a[0], a[1], a[2], dv[0], dv[1], dv[2]);
  LAUNCHCHECK("k_kick");
  for (int q = 0; q < 3; ++q)
    HIPCHECK(hipMemcpyAsync(hv[q], dv[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int tgrav_set_bidir_config(tgrav_ctx* c, int mode, int variant, int nsplit) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc = check_mode(mode);
  if (rc) return rc;
  if (variant >= kNumBi) return fail(TGRAV_EINVAL, "no fused variant %d", variant);
  if (variant >= 0) c->bvariant[mode] = variant;
  if (nsplit >= 0) c->bnsplit[mode] = nsplit;
  return 0;
}

int tgrav_bidir_variant_info(int mode, int variant, int* k, int* threads) {
  int rc = check_mode(mode);
  if (rc) return rc;
  if (variant < 0 || variant >= kNumBi) return fail(TGRAV_EINVAL, "no fused variant %d", variant);
  if (k) *k = kBiVariants[mode][variant].k;
  if (threads) *threads = kBiVariants[mode][variant].nt;
  return 0;
}

// Fused passes: packs the particles (x, y, z, m, eps) into c->bp4 / c->bpe2 and their smallest
// eps^2 into c->bstat.
static int bidir_pack_particles(tgrav_ctx* c, int64_t np, const double* const* pa) {
  int rc;
  if ((rc = c->bp4.ensure((size_t)np * sizeof(double4))) ||
      (rc = c->bpe2.ensure((size_t)np * sizeof(double))) ||
      (rc = c->bstat.ensure(tg::ST_WORDS * sizeof(unsigned long long))))
    return rc;
  tg::k_reset_stats<<<1, 64, 0, c->stream>>>(c->bstat.as<unsigned long long>());
  LAUNCHCHECK("k_reset_stats");
  tg::k_pack_sources<<<grid1(np, 256), 256, 0, c->stream>>>(
      np, pa[0], pa[1], pa[2], pa[3], pa[4], c->bp4.as<double4>(), c->bpe2.as<double>(),
      c->bstat.as<unsigned long long>(), 0);
  LAUNCHCHECK("k_pack_sources");
  return 0;
}

// Launch shape of the block-column fused kernel: particle chunks over blockIdx.y, about 8 rounds
// of one-wave workgroups, at most 12 chunks, at least 4 particle tiles a chunk.
static int bidir_grid_shape(const tgrav_ctx* c, int mode, int64_t nblk, int64_t np, int64_t* chunk,
                            int* nsplit) {
  int64_t want = c->bnsplit[mode];
  if (want <= 0) {
    const int64_t target = 8LL * 12 * c->num_cu;
    want = std::min<int64_t>((target + nblk - 1) / nblk,
                             std::min<int64_t>(12, std::max<int64_t>(1, np / 256)));
  }
  want = std::max<int64_t>(1, std::min<int64_t>(want, 65535));
  *chunk = (np + want - 1) / want;
  *chunk = (*chunk + 63) / 64 * 64;
  *nsplit = (int)((np + *chunk - 1) / *chunk);
  const double part_bytes = (double)nblk * 4 * np * 8 + (*nsplit > 1 ? (double)*nsplit * 4 * nblk * 512 * 8 : 0.0);
  if (part_bytes > 64.0 * (1ULL << 30))
    return fail(TGRAV_ELIMIT, "fused pass would need %.1f GiB of partial sums; evaluate the two "
                "directions separately", part_bytes / (1ULL << 30));
  return 0;
}

static void bidir_grid_kernel(tgrav_ctx* c, int mode, const tg::CellBlocks& C, int64_t np, int64_t chunk,
                              int nsplit, const tg::Out& o) {
  const dim3 g((unsigned)C.nblk, (unsigned)nsplit);
  if (mode == TGRAV_MIXED)
    tg::k_bidir_grid<true><<<g, 64, 0, c->stream>>>(C, c->bp4.as<double4>(), c->bpe2.as<double>(), np,
                                                   chunk, c->bstat.as<const double>(), o,
                                                   c->bppart.as<double>());
  else
    tg::k_bidir_grid<false><<<g, 64, 0, c->stream>>>(C, c->bp4.as<double4>(), c->bpe2.as<double>(), np,
                                                    chunk, c->bstat.as<const double>(), o,
                                                    c->bppart.as<double>());
}

// The fused pass with the cells as 8^3 block descriptors (k_bidir_grid).
static int bidir_grid_pass(tgrav_ctx* c, int mode, const tg::CellBlocks& C, int64_t np,
                           const double* const* pa, const tg::Out& CO, const tg::Out& PO) {
  int rc;
  const int64_t nblk = C.nblk, nc = nblk * 512;
  if ((rc = bidir_pack_particles(c, np, pa))) return rc;
  int64_t chunk;
  int nsplit;
  if ((rc = bidir_grid_shape(c, mode, nblk, np, &chunk, &nsplit))) return rc;
  if ((rc = c->bppart.ensure((size_t)nblk * 4 * np * sizeof(double)))) return rc;
  tg::Out o = CO;
  if (nsplit > 1) {
    if ((rc = c->part.ensure((size_t)nsplit * 4 * nc * sizeof(double)))) return rc;
    o.part = c->part.as<double>();
    o.pstride = nc;
  }
  rc = timed_launch(c, nc, np, [&]() -> int {
    bidir_grid_kernel(c, mode, C, np, chunk, nsplit, o);
    LAUNCHCHECK("k_bidir_grid");
    return 0;
  });
  if (rc) return rc;
  if (nsplit > 1) {
    if ((rc = reduce_chunks(c, c->part.as<double>(), nc, nsplit, nc, CO))) return rc;
  }
  return reduce_particle_partials(c, nblk, np, PO);
}

int tgrav_bidirectional(tgrav_ctx* c, int64_t nc, const double* cx, const double* cy,
                        const double* cz, const double* cm, const double* ce, int64_t np,
                        const double* px, const double* py, const double* pz, const double* pm,
                        const double* pe, double* cax, double* cay, double* caz, double* cphi,
                        double* pax, double* pay, double* paz, double* pphi, int mode, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc;
  if ((rc = check_mode(mode)) || (rc = check_kind(kind))) return rc;
  if (c->soft != TGRAV_SOFT_PLUMMER || c->periodic)
    return fail(TGRAV_EINVAL, "the fused pass uses the Plummer kernel (tgrav_set_softening(ctx, 0, 0, 0))");
  if (nc < 0 || np < 0) return fail(TGRAV_EINVAL, "sizes must be >= 0");
  if (nc > 0 && (!cx || !cy || !cz || !cm)) return fail(TGRAV_EINVAL, "cell x, y, z and m are required");
  if (np > 0 && (!px || !py || !pz || !pm))
    return fail(TGRAV_EINVAL, "particle x, y, z and m are required");
  if ((rc = set_device(c))) return rc;
  const double* ca[5] = {cx, cy, cz, cm, ce};
  const double* pa[5] = {px, py, pz, pm, pe};
  double* co[4] = {cax, cay, caz, cphi};
  double* po[4] = {pax, pay, paz, pphi};
  if (kind == TGRAV_HOST && c->validate) {
    for (int q = 0; q < 5; ++q)
      if (!host_finite(nc, ca[q]) || !host_finite(np, pa[q]))
        return fail(TGRAV_ENONFINITE, "non-finite value in an input array");
  }
  if (kind == TGRAV_HOST) {
    if ((rc = stage_in(c, 0, nc, ca, 5)) || (rc = stage_in(c, 5, np, pa, 5))) return rc;
    for (int q = 0; q < 4; ++q) {
      if (co[q]) {
        if ((rc = c->outb[q].ensure((size_t)std::max<int64_t>(nc, 1) * 8))) return rc;
      }
      if (po[q]) {
        if ((rc = c->outb[4 + q].ensure((size_t)std::max<int64_t>(np, 1) * 8))) return rc;
      }
    }
  } else if (c->validate) {
    bool bad = false;
    if ((rc = device_check_finite(c, nc, ca, 5, &bad))) return rc;
    if (!bad && (rc = device_check_finite(c, np, pa, 5, &bad))) return rc;
    if (bad) return fail(TGRAV_ENONFINITE, "non-finite value in an input array");
  }
  if (mode == TGRAV_MIXED && c->validate && nc > 0 && np > 0) {
    if ((rc = range_stats(c, 0, nc, ca[0], ca[1], ca[2], nullptr, ca[4], nc, ca[3])) ||
        (rc = range_stats(c, 1, np, pa[0], pa[1], pa[2], nullptr, pa[4], np, pa[3])) ||
        (rc = mixed_range_ok(host_stats(c, 0), host_stats(c, 1), "the cells and particles")) ||
        (rc = mixed_range_ok(host_stats(c, 1), host_stats(c, 0), "the particles and cells")))
      return rc;
  }
  double* dco[4];
  double* dpo[4];
  for (int q = 0; q < 4; ++q) {
    dco[q] = kind == TGRAV_HOST ? (co[q] ? c->outb[q].as<double>() : nullptr) : co[q];
    dpo[q] = kind == TGRAV_HOST ? (po[q] ? c->outb[4 + q].as<double>() : nullptr) : po[q];
  }
  tg::Out CO{dco[0], dco[1], dco[2], dco[3], nullptr, nullptr, nullptr, 0.0, nullptr, 0, c->G};
  tg::Out PO{dpo[0], dpo[1], dpo[2], dpo[3], nullptr, nullptr, nullptr, 0.0, nullptr, 0, c->G};
  if (nc > 0 && np == 0) {
    tg::k_fill<<<grid1(nc, 256), 256, 0, c->stream>>>(nc, 0.0, dco[0], dco[1], dco[2], dco[3]);
    LAUNCHCHECK("k_fill");
  }
  if (np > 0 && nc == 0) {
    tg::k_fill<<<grid1(np, 256), 256, 0, c->stream>>>(np, 0.0, dpo[0], dpo[1], dpo[2], dpo[3]);
    LAUNCHCHECK("k_fill");
  }
  if (nc > 0 && np > 0) {
    if ((rc = bidir_pack_particles(c, np, pa))) return rc;
    // Cells in the C order of 8^3 leaf-block meshgrids take the block-column kernel.  Whether
    // they are is tested on the device (k_check_cols writes a gate word: 1 columns, 0 not); both
    // kernels and their reductions are launched, and the word lets one of them run.
    bool cand = nc % 512 == 0 && c->block_kernel;
    const BiVariant& v = kBiVariants[mode][c->bvariant[mode]];
    const int64_t tile = (int64_t)v.k * v.nt;
    const int64_t ntiles = (nc + tile - 1) / tile;
    if (ntiles > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many cells");
    int want = c->bnsplit[mode];
    if (want <= 0) {
      const int64_t target_wg = 4LL * c->num_cu;
      want = ntiles >= target_wg / 2 ? 1
                                     : (int)std::min<int64_t>((target_wg + ntiles - 1) / ntiles,
                                                              std::max<int64_t>(1, np / (4LL * v.nt)));
    }
    if (want > 65535) want = 65535;
    int64_t chunk = (np + want - 1) / want;
    chunk = (chunk + v.nt - 1) / v.nt * v.nt;
    const int nsplit = (int)((np + chunk - 1) / chunk);
    const double part_bytes = (double)ntiles * 4 * np * 8 + (nsplit > 1 ? (double)nsplit * 4 * nc * 8 : 0.0);
    if (part_bytes > 64.0 * (1ULL << 30))
      return fail(TGRAV_ELIMIT, "fused pass would need %.1f GiB of partial sums; evaluate the two "
                  "directions separately", part_bytes / (1ULL << 30));
    const int64_t nblk = nc / 512;
    int64_t chunk_g = 0;
    int nsplit_g = 1;
    // The size limit is the generic pipeline's; a column pipeline that would exceed it is simply
    // not offered (the generic one then runs whatever the structure).
    if (cand && bidir_grid_shape(c, mode, nblk, np, &chunk_g, &nsplit_g) != 0) cand = false;
    if (cand && ((double)std::max<int64_t>(ntiles, nblk) * 4 * np * 8 +
                 (std::max(nsplit, nsplit_g) > 1 ? (double)std::max(nsplit, nsplit_g) * 4 * nc * 8 : 0.0)) >
                    64.0 * (1ULL << 30))
      cand = false;
    const int64_t prows = std::max<int64_t>(ntiles, cand ? nblk : 0);
    if ((rc = c->bppart.ensure((size_t)prows * 4 * np * sizeof(double)))) return rc;
    const int nsp = std::max(nsplit, cand ? nsplit_g : 1);
    if (nsp > 1 && (rc = c->part.ensure((size_t)nsp * 4 * nc * sizeof(double)))) return rc;
    tg::Out CG = CO, PG = PO, CX = CO, PX = PO;   // column pipeline (G), generic pipeline (X)
    if (cand) {
      if ((rc = c->colflag.ensure(sizeof(int)))) return rc;
      int* fl = c->colflag.as<int>();
      tg::Tgt T{};
      T.x = ca[0]; T.y = ca[1]; T.z = ca[2]; T.eps = ca[4]; T.n = nc;
      tg::k_set_flag<<<1, 64, 0, c->stream>>>(fl, 1);
      LAUNCHCHECK("k_set_flag");
      tg::k_check_cols<<<grid1(nc, 256), 256, 0, c->stream>>>(T, fl);
      LAUNCHCHECK("k_check_cols");
      CG.gate = PG.gate = CX.gate = PX.gate = fl;
      CG.gate_on = PG.gate_on = 1;
      CX.gate_on = PX.gate_on = 0;
    }
    tg::Out og = CG, ox = CX;
    og.part = nsplit_g > 1 ? c->part.as<double>() : nullptr;
    ox.part = nsplit > 1 ? c->part.as<double>() : nullptr;
    og.pstride = ox.pstride = nc;
    const tg::CellBlocks CB{nullptr, nullptr, nullptr, nullptr, ca[3], nullptr, nblk, ca[0], ca[1], ca[2], ca[4]};
    tg::Cells C{ca[0], ca[1], ca[2], ca[3], ca[4], nc};
    rc = timed_launch(c, nc, np, [&]() -> int {
      if (cand) {
        bidir_grid_kernel(c, mode, CB, np, chunk_g, nsplit_g, og);
        LAUNCHCHECK("k_bidir_grid");
      }
      v.fn(dim3((unsigned)ntiles, (unsigned)nsplit), dim3(v.nt), c->stream, C, c->bp4.as<double4>(),
           c->bpe2.as<double>(), np, chunk, c->bstat.as<const double>(), ox, c->bppart.as<double>());
      LAUNCHCHECK("k_bidir");
      return 0;
    });
    if (rc) return rc;
    if (cand) {
      if (nsplit_g > 1 && (rc = reduce_chunks(c, c->part.as<double>(), nc, nsplit_g, nc, CG))) return rc;
      if ((rc = reduce_particle_partials(c, nblk, np, PG))) return rc;
    }
    if (nsplit > 1 && (rc = reduce_chunks(c, c->part.as<double>(), nc, nsplit, nc, CX))) return rc;
    if ((rc = reduce_particle_partials(c, ntiles, np, PX))) return rc;
  }
  if (kind == TGRAV_HOST) {
    for (int q = 0; q < 4; ++q) {
      if (co[q] && nc > 0)
        HIPCHECK(hipMemcpyAsync(co[q], dco[q], (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
      if (po[q] && np > 0)
        HIPCHECK(hipMemcpyAsync(po[q], dpo[q], (size_t)np * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int tgrav_bidirectional_blocks(tgrav_ctx* c, int64_t nblk, int nb, const double* ox,
                               const double* oy, const double* oz, const double* dx,
                               const double* cm, const double* ceps, int64_t np,
                               const double* px, const double* py, const double* pz,
                               const double* pm, const double* pe, double* cax, double* cay,
                               double* caz, double* cphi, double* pax, double* pay, double* paz,
                               double* pphi, int mode, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc;
  if ((rc = check_mode(mode)) || (rc = check_kind(kind))) return rc;
  if (c->soft != TGRAV_SOFT_PLUMMER || c->periodic)
    return fail(TGRAV_EINVAL, "the fused pass uses the Plummer kernel (tgrav_set_softening(ctx, 0, 0, 0))");
  if (nb != 8) return fail(TGRAV_EINVAL, "cell blocks must be 8^3 (nb = 8), got %d", nb);
  if (nblk < 0 || np < 0) return fail(TGRAV_EINVAL, "sizes must be >= 0");
  if (nblk > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many blocks");
  const int64_t nc = nblk * 512;
  if (nblk > 0 && (!ox || !oy || !oz || !dx || !cm))
    return fail(TGRAV_EINVAL, "block corners, widths and cell masses are required");
  if (np > 0 && (!px || !py || !pz || !pm))
    return fail(TGRAV_EINVAL, "particle x, y, z and m are required");
  if ((rc = set_device(c))) return rc;
  const double* ba[4] = {ox, oy, oz, dx};
  const double* mc[1] = {cm};
  const double* be[1] = {ceps};
  const double* pa[5] = {px, py, pz, pm, pe};
  double* co[4] = {cax, cay, caz, cphi};
  double* po[4] = {pax, pay, paz, pphi};
  if (kind == TGRAV_HOST && c->validate) {
    bool ok = host_finite(nc, cm) && host_finite(nblk, ceps);
    for (int q = 0; q < 4; ++q) ok = ok && host_finite(nblk, ba[q]);
    for (int q = 0; q < 5; ++q) ok = ok && host_finite(np, pa[q]);
    if (!ok) return fail(TGRAV_ENONFINITE, "non-finite value in an input array");
  }
  if (kind == TGRAV_HOST) {
    if ((rc = stage_in(c, 0, nblk, ba, 4)) || (rc = stage_in(c, 4, nc, mc, 1)) ||
        (rc = stage_in(c, 10, nblk, be, 1)) || (rc = stage_in(c, 5, np, pa, 5)))
      return rc;
    for (int q = 0; q < 4; ++q) {
      if (co[q] && (rc = c->outb[q].ensure((size_t)std::max<int64_t>(nc, 1) * 8))) return rc;
      if (po[q] && (rc = c->outb[4 + q].ensure((size_t)std::max<int64_t>(np, 1) * 8))) return rc;
    }
  } else if (c->validate) {
    bool bad = false;
    if ((rc = device_check_finite(c, nblk, ba, 4, &bad))) return rc;
    if (!bad && (rc = device_check_finite(c, nc, mc, 1, &bad))) return rc;
    if (!bad && (rc = device_check_finite(c, nblk, be, 1, &bad))) return rc;
    if (!bad && (rc = device_check_finite(c, np, pa, 5, &bad))) return rc;
    if (bad) return fail(TGRAV_ENONFINITE, "non-finite value in an input array");
  }
  if (mode == TGRAV_MIXED && c->validate && nc > 0 && np > 0) {
    if ((rc = range_stats(c, 0, nblk, ba[0], ba[1], ba[2], ba[3], be[0], nc, mc[0])) ||
        (rc = range_stats(c, 1, np, pa[0], pa[1], pa[2], nullptr, pa[4], np, pa[3])) ||
        (rc = mixed_range_ok(host_stats(c, 0), host_stats(c, 1), "the cells and particles")) ||
        (rc = mixed_range_ok(host_stats(c, 1), host_stats(c, 0), "the particles and cells")))
      return rc;
  }
  double* dco[4];
  double* dpo[4];
  for (int q = 0; q < 4; ++q) {
    dco[q] = kind == TGRAV_HOST ? (co[q] ? c->outb[q].as<double>() : nullptr) : co[q];
    dpo[q] = kind == TGRAV_HOST ? (po[q] ? c->outb[4 + q].as<double>() : nullptr) : po[q];
  }
  tg::Out CO{dco[0], dco[1], dco[2], dco[3], nullptr, nullptr, nullptr, 0.0, nullptr, 0, c->G};
  tg::Out PO{dpo[0], dpo[1], dpo[2], dpo[3], nullptr, nullptr, nullptr, 0.0, nullptr, 0, c->G};
  if (nc > 0 && np == 0) {
    tg::k_fill<<<grid1(nc, 256), 256, 0, c->stream>>>(nc, 0.0, dco[0], dco[1], dco[2], dco[3]);
    LAUNCHCHECK("k_fill");
  }
  if (np > 0 && nc == 0) {
    tg::k_fill<<<grid1(np, 256), 256, 0, c->stream>>>(np, 0.0, dpo[0], dpo[1], dpo[2], dpo[3]);
    LAUNCHCHECK("k_fill");
  }
  if (nc > 0 && np > 0) {
    tg::CellBlocks C{ba[0], ba[1], ba[2], ba[3], mc[0], be[0], nblk, nullptr, nullptr, nullptr, nullptr};
    if ((rc = bidir_grid_pass(c, mode, C, np, pa, CO, PO))) return rc;
  }
  if (kind == TGRAV_HOST) {
    for (int q = 0; q < 4; ++q) {
      if (co[q] && nc > 0)
        HIPCHECK(hipMemcpyAsync(co[q], dco[q], (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
      if (po[q] && np > 0)
        HIPCHECK(hipMemcpyAsync(po[q], dpo[q], (size_t)np * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int tgrav_set_softening(tgrav_ctx* c, int kernel, double rsoft, int flags) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  if (kernel != TGRAV_SOFT_PLUMMER && kernel != TGRAV_SOFT_FLASH_SPLINE &&
      kernel != TGRAV_SOFT_FLASH_LINEAR)
    return fail(TGRAV_EINVAL, "unknown softening kernel %d", kernel);
  if (flags & ~TGRAV_FLASH_QUIRKS) return fail(TGRAV_EINVAL, "unknown softening flags 0x%x", flags);
  if (kernel != TGRAV_SOFT_PLUMMER && !(rsoft > 0.0 && rsoft < 1e150))
    return fail(TGRAV_EINVAL, "the softening radius must be positive and finite");
  c->soft = kernel;
  c->rsoft = kernel == TGRAV_SOFT_PLUMMER ? 0.0 : rsoft;
  c->soft_quirks = (flags & TGRAV_FLASH_QUIRKS) ? 1 : 0;
  return 0;
}

int tgrav_set_periodic(tgrav_ctx* c, int on, double lx, double ly, double lz, int nrep) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  if (!on) {
    c->periodic = 0;
    return 0;
  }
  if (!(lx > 0.0 && ly > 0.0 && lz > 0.0 && lx < 1e150 && ly < 1e150 && lz < 1e150))
    return fail(TGRAV_EINVAL, "box lengths must be positive and finite");
  if (nrep < 1 || nrep > 64) return fail(TGRAV_EINVAL, "nrep must be in [1, 64], got %d", nrep);
  c->periodic = on == 2 ? 2 : 1;
  c->box[0] = lx;
  c->box[1] = ly;
  c->box[2] = lz;
  c->nrep = nrep;
  return 0;
}

// ------------------------------------------------------------ particle mesh ------
// Host pointers are staged into c->pm[slot]; device pointers pass through.
static int pm_in(tgrav_ctx* c, int kind, int slot, const void* h, size_t bytes, const void** out) {
  *out = h;
  if (kind == TGRAV_DEVICE || !h || bytes == 0) return 0;
  int rc = c->pm[slot].ensure(bytes);
  if (rc) return rc;
  HIPCHECK(hipMemcpyAsync(c->pm[slot].p, h, bytes, hipMemcpyHostToDevice, c->stream));
  *out = c->pm[slot].p;
  return 0;
}

static int pm_out(tgrav_ctx* c, int kind, int slot, void* h, size_t bytes, bool upload,
                  void** out) {
  *out = h;
  if (kind == TGRAV_DEVICE || !h || bytes == 0) return 0;
  int rc = c->pm[slot].ensure(bytes);
  if (rc) return rc;
  if (upload) HIPCHECK(hipMemcpyAsync(c->pm[slot].p, h, bytes, hipMemcpyHostToDevice, c->stream));
  *out = c->pm[slot].p;
  return 0;
}

static int pm_back(tgrav_ctx* c, int kind, void* h, const void* d, size_t bytes) {
  if (kind == TGRAV_DEVICE || !h || bytes == 0) return 0;
  HIPCHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
  return 0;
}

static int pm_geometry(int64_t nblk, int nb, int ng, int64_t* ncell) {
  if (nblk < 0) return fail(TGRAV_EINVAL, "nblk must be >= 0");
  if (nb < 1 || ng < 1 || nb + 2 * ng > 128)
    return fail(TGRAV_EINVAL, "need nb >= 1, ng >= 1 and nb + 2 ng <= 128 (got nb %d, ng %d)", nb, ng);
  const int64_t nbg = nb + 2 * ng;
  if (nblk > ((int64_t)1 << 40) / (nbg * nbg * nbg)) return fail(TGRAV_ELIMIT, "too many cells");
  *ncell = nblk * nbg * nbg * nbg;
  return 0;
}

static int pm_read_err(tgrav_ctx* c, const char* what) {
  bool bad = false;
  int rc = read_flag(c, c->pm_stat.as<unsigned long long>() + 1, &bad);
  if (rc) return rc;
  if (bad)
    return fail(TGRAV_EINVAL, "%s: a particle's block index is out of range or its cloud leaves "
                "its guarded block", what);
  return 0;
}

int tgrav_pm_deposit(tgrav_ctx* c, int64_t np, const double* x, const double* y, const double* z,
                     const double* m, const int64_t* blk, int64_t nblk, int nb, int ng,
                     const double* ox, const double* oy, const double* oz, const double* bsize,
                     double* rho, int accumulate, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc;
  int64_t ncell;
  if ((rc = check_kind(kind)) || (rc = pm_geometry(nblk, nb, ng, &ncell))) return rc;
  if (np < 0) return fail(TGRAV_EINVAL, "np must be >= 0");
  if (np > ((int64_t)1 << 28)) return fail(TGRAV_ELIMIT, "at most 2^28 particles per deposit");
  if (np > 0 && (!x || !y || !z || !m || !blk)) return fail(TGRAV_EINVAL, "x, y, z, m and blk are required");
  if (nblk > 0 && (!ox || !oy || !oz || !bsize || !rho))
    return fail(TGRAV_EINVAL, "ox, oy, oz, bsize and rho are required");
  if (np > 0 && nblk == 0) return fail(TGRAV_EINVAL, "particles need blocks");
  if ((rc = set_device(c))) return rc;
  if (kind == TGRAV_HOST && c->validate) {
    const double* arr[8] = {x, y, z, m, ox, oy, oz, bsize};
    for (int q = 0; q < 8; ++q)
      if (!host_finite(q < 4 ? np : nblk, arr[q])) return fail(TGRAV_ENONFINITE, "non-finite input %d", q);
  }
  const void *dx_, *dy_, *dz_, *dm_, *db_, *dox, *doy, *doz, *dbs;
  void* drho;
  if ((rc = pm_in(c, kind, 0, x, np * 8, &dx_)) || (rc = pm_in(c, kind, 1, y, np * 8, &dy_)) ||
      (rc = pm_in(c, kind, 2, z, np * 8, &dz_)) || (rc = pm_in(c, kind, 3, m, np * 8, &dm_)) ||
      (rc = pm_in(c, kind, 4, blk, np * 8, &db_)) || (rc = pm_in(c, kind, 5, ox, nblk * 8, &dox)) ||
      (rc = pm_in(c, kind, 6, oy, nblk * 8, &doy)) || (rc = pm_in(c, kind, 7, oz, nblk * 8, &doz)) ||
      (rc = pm_in(c, kind, 8, bsize, nblk * 8, &dbs)) ||
      (rc = pm_out(c, kind, 9, rho, ncell * 8, accumulate != 0, &drho)))
    return rc;
  if (ncell == 0) return 0;
  if ((rc = c->pm_acc[0].ensure(ncell * 8)) || (rc = c->pm_acc[1].ensure(ncell * 8)) ||
      (rc = c->pm_acc[2].ensure(ncell * 8)) || (rc = c->pm_stat.ensure(16)))
    return rc;
  tg::PmBlocks B{(const double*)dox, (const double*)doy, (const double*)doz, (const double*)dbs,
                 nblk, nb, ng, nb + 2 * ng, (int64_t)(nb + 2 * ng) * (nb + 2 * ng) * (nb + 2 * ng)};
  unsigned long long* H = c->pm_acc[0].as<unsigned long long>();
  unsigned long long* L = c->pm_acc[1].as<unsigned long long>();
  unsigned long long* M = c->pm_acc[2].as<unsigned long long>();
  unsigned long long* st = c->pm_stat.as<unsigned long long>();
  tg::k_zero_u64<<<grid1(ncell, 256), 256, 0, c->stream>>>(ncell, H, L, M);
  tg::k_zero_u64<<<1, 64, 0, c->stream>>>(2, st, nullptr, nullptr);
  LAUNCHCHECK("k_zero_u64");
  if (np > 0) {
    tg::k_pm_cell_max<<<grid1(np, 256), 256, 0, c->stream>>>(
        np, (const double*)dx_, (const double*)dy_, (const double*)dz_, (const double*)dm_,
        (const int64_t*)db_, B, M, st + 1);
    LAUNCHCHECK("k_pm_cell_max");
    if (c->validate && (rc = pm_read_err(c, "tgrav_pm_deposit"))) return rc;
    rc = timed_launch(c, np, ncell, [&]() -> int {
      tg::k_pm_deposit<<<grid1(np, 256), 256, 0, c->stream>>>(
          np, (const double*)dx_, (const double*)dy_, (const double*)dz_, (const double*)dm_,
          (const int64_t*)db_, B, M, H, L);
      LAUNCHCHECK("k_pm_deposit");
      return 0;
    });
    if (rc) return rc;
  }
  tg::k_pm_finalize<<<grid1(ncell, 256), 256, 0, c->stream>>>(ncell, H, L, M, np, (double*)drho,
                                                               accumulate != 0);
  LAUNCHCHECK("k_pm_finalize");
  if ((rc = pm_back(c, kind, rho, drho, ncell * 8))) return rc;
  if (kind == TGRAV_HOST) HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int tgrav_pm_gradient(tgrav_ctx* c, int64_t nblk, int nb, int ng, const double* bsize,
                      const double* phi, double* gx, double* gy, double* gz, double dt,
                      double* vx, double* vy, double* vz, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc;
  int64_t ncell;
  if ((rc = check_kind(kind)) || (rc = pm_geometry(nblk, nb, ng, &ncell))) return rc;
  if (nblk > 0 && (!bsize || !phi)) return fail(TGRAV_EINVAL, "bsize and phi are required");
  if ((vx || vy || vz) && !(vx && vy && vz)) return fail(TGRAV_EINVAL, "vx, vy, vz go together");
  if (!isfinite(dt)) return fail(TGRAV_EINVAL, "dt must be finite");
  if ((rc = set_device(c))) return rc;
  if (kind == TGRAV_HOST && c->validate &&
      (!host_finite(nblk, bsize) || !host_finite(ncell, phi)))
    return fail(TGRAV_ENONFINITE, "non-finite block size or potential");
  const void *dbs, *dphi;
  void *dg[3], *dv[3];
  if ((rc = pm_in(c, kind, 0, bsize, nblk * 8, &dbs)) || (rc = pm_in(c, kind, 1, phi, ncell * 8, &dphi)) ||
      (rc = pm_out(c, kind, 2, gx, ncell * 8, false, &dg[0])) ||
      (rc = pm_out(c, kind, 3, gy, ncell * 8, false, &dg[1])) ||
      (rc = pm_out(c, kind, 4, gz, ncell * 8, false, &dg[2])) ||
      (rc = pm_out(c, kind, 5, vx, ncell * 8, true, &dv[0])) ||
      (rc = pm_out(c, kind, 6, vy, ncell * 8, true, &dv[1])) ||
      (rc = pm_out(c, kind, 7, vz, ncell * 8, true, &dv[2])))
    return rc;
  if (ncell == 0) return 0;
  tg::PmBlocks B{nullptr, nullptr, nullptr, (const double*)dbs, nblk, nb, ng, nb + 2 * ng,
                 (int64_t)(nb + 2 * ng) * (nb + 2 * ng) * (nb + 2 * ng)};
  const uint32_t nbg = (uint32_t)(nb + 2 * ng);
  const uint32_t magic = (uint32_t)((((uint64_t)1 << 32) + nbg - 1) / nbg);   // ceil(2^32 / nbg)
  const unsigned S = (unsigned)((B.nc + 255) / 256);                       // slices per block
  const int64_t ngrp = (nblk + 7) / 8;                                     // groups of 8 blocks
  rc = timed_launch(c, ncell, 0, [&]() -> int {
    for (int64_t g0 = 0; g0 < ngrp; g0 += 65535) {
      const dim3 g(8 * S, (unsigned)std::min<int64_t>(65535, ngrp - g0));
      tg::k_pm_gradient<<<g, 256, 0, c->stream>>>(
          B, magic, g0, (const double*)dphi, (double*)dg[0], (double*)dg[1], (double*)dg[2], dt,
          (double*)dv[0], (double*)dv[1], (double*)dv[2]);
    }
    LAUNCHCHECK("k_pm_gradient");
    return 0;
  });
  if (rc) return rc;
  double* hs[6] = {gx, gy, gz, vx, vy, vz};
  void* ds[6] = {dg[0], dg[1], dg[2], dv[0], dv[1], dv[2]};
  for (int q = 0; q < 6; ++q)
    if ((rc = pm_back(c, kind, hs[q], ds[q], ncell * 8))) return rc;
  if (kind == TGRAV_HOST) HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int tgrav_pm_interpolate(tgrav_ctx* c, int64_t np, const double* x, const double* y,
                         const double* z, const int64_t* blk, int64_t nblk, int nb, int ng,
                         const double* ox, const double* oy, const double* oz,
                         const double* bsize, const double* gx, const double* gy,
                         const double* gz, double* ax, double* ay, double* az, int kind) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  int rc;
  int64_t ncell;
  if ((rc = check_kind(kind)) || (rc = pm_geometry(nblk, nb, ng, &ncell))) return rc;
  if (np < 0) return fail(TGRAV_EINVAL, "np must be >= 0");
  if (np > 0 && (!x || !y || !z || !blk || !ax || !ay || !az))
    return fail(TGRAV_EINVAL, "x, y, z, blk and ax, ay, az are required");
  if (np > 0 && (nblk == 0 || !ox || !oy || !oz || !bsize || !gx || !gy || !gz))
    return fail(TGRAV_EINVAL, "blocks (ox, oy, oz, bsize) and fields gx, gy, gz are required");
  if ((rc = set_device(c))) return rc;
  if (np == 0) return 0;
  if (kind == TGRAV_HOST && c->validate) {
    const double* arr[7] = {x, y, z, ox, oy, oz, bsize};
    for (int q = 0; q < 7; ++q)
      if (!host_finite(q < 3 ? np : nblk, arr[q])) return fail(TGRAV_ENONFINITE, "non-finite input %d", q);
  }
  const void *dx_, *dy_, *dz_, *db_, *dox, *doy, *doz, *dbs, *dgx, *dgy, *dgz;
  void *dax, *day, *daz;
  if ((rc = pm_in(c, kind, 0, x, np * 8, &dx_)) || (rc = pm_in(c, kind, 1, y, np * 8, &dy_)) ||
      (rc = pm_in(c, kind, 2, z, np * 8, &dz_)) || (rc = pm_in(c, kind, 3, blk, np * 8, &db_)) ||
      (rc = pm_in(c, kind, 4, ox, nblk * 8, &dox)) || (rc = pm_in(c, kind, 5, oy, nblk * 8, &doy)) ||
      (rc = pm_in(c, kind, 6, oz, nblk * 8, &doz)) || (rc = pm_in(c, kind, 7, bsize, nblk * 8, &dbs)) ||
      (rc = pm_in(c, kind, 8, gx, ncell * 8, &dgx)) || (rc = pm_in(c, kind, 9, gy, ncell * 8, &dgy)) ||
      (rc = pm_in(c, kind, 10, gz, ncell * 8, &dgz)) ||
      (rc = pm_out(c, kind, 11, ax, np * 8, false, &dax)) ||
      (rc = pm_out(c, kind, 12, ay, np * 8, false, &day)) ||
      (rc = pm_out(c, kind, 13, az, np * 8, false, &daz)))
    return rc;
  if ((rc = c->pm_stat.ensure(16))) return rc;
  unsigned long long* st = c->pm_stat.as<unsigned long long>();
  tg::k_zero_u64<<<1, 64, 0, c->stream>>>(2, st, nullptr, nullptr);
  LAUNCHCHECK("k_zero_u64");
  tg::PmBlocks B{(const double*)dox, (const double*)doy, (const double*)doz, (const double*)dbs,
                 nblk, nb, ng, nb + 2 * ng, (int64_t)(nb + 2 * ng) * (nb + 2 * ng) * (nb + 2 * ng)};
  rc = timed_launch(c, np, ncell, [&]() -> int {
    tg::k_pm_interpolate<<<grid1(np, 256), 256, 0, c->stream>>>(
        np, (const double*)dx_, (const double*)dy_, (const double*)dz_, (const int64_t*)db_, B,
        (const double*)dgx, (const double*)dgy, (const double*)dgz, (double*)dax, (double*)day,
        (double*)daz, st + 1);
    LAUNCHCHECK("k_pm_interpolate");
    return 0;
  });
  if (rc) return rc;
  if (c->validate && (rc = pm_read_err(c, "tgrav_pm_interpolate"))) return rc;
  if ((rc = pm_back(c, kind, ax, dax, np * 8)) || (rc = pm_back(c, kind, ay, day, np * 8)) ||
      (rc = pm_back(c, kind, az, daz, np * 8)))
    return rc;
  if (kind == TGRAV_HOST) HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int tgrav_set_block_kernel(tgrav_ctx* c, int on) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  c->block_kernel = on ? 1 : 0;
  return 0;
}

int tgrav_set_column_waves(tgrav_ctx* c, int waves) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  if (waves != 1 && waves != 2 && waves != 4)
    return fail(TGRAV_EINVAL, "waves per column workgroup must be 1, 2 or 4 (got %d)", waves);
  c->col_waves = waves;
  return 0;
}

int tgrav_get_column_waves(tgrav_ctx* c) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  return c->col_waves;
}

int tgrav_set_timing(tgrav_ctx* c, int on) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  c->timing = on ? 1 : 0;
  c->ev_used = 0;
  return 0;
}

int tgrav_timing_read(tgrav_ctx* c, int max, double* ms, int64_t* nt, int64_t* ns) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  if (max < 0 || (max > 0 && !ms)) return fail(TGRAV_EINVAL, "bad output buffer");
  int rc = set_device(c);
  if (rc) return rc;
  const int n = std::min(max, c->ev_used);
  for (int k = 0; k < n; ++k) {
    HIPCHECK(hipEventSynchronize(c->ev[2 * k + 1]));
    float f = 0.f;
    HIPCHECK(hipEventElapsedTime(&f, c->ev[2 * k], c->ev[2 * k + 1]));
    ms[k] = f;
    if (nt) nt[k] = c->ev_nt[k];
    if (ns) ns[k] = c->ev_ns[k];
  }
  const int total = c->ev_used;
  c->ev_used = 0;
  return total;
}

int tgrav_set_batch_config(tgrav_ctx* c, double item_interactions, int wg_per_cu) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  if (!(item_interactions >= 0.0) || wg_per_cu < 0)
    return fail(TGRAV_EINVAL, "item size and workgroups per CU must be >= 0");
  c->batch_item = item_interactions;
  c->batch_wg_per_cu = wg_per_cu;
  c->user_batch = true;
  return 0;
}

int tgrav_set_batch_scales(tgrav_ctx* c, double column_scale, double block_source_scale) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  if (!(column_scale >= 0.0 && column_scale <= 64.0 && block_source_scale >= 0.0 && block_source_scale <= 64.0))
    return fail(TGRAV_EINVAL, "item scales must be in [0, 64]");
  const bool reset = column_scale == 0.0 && block_source_scale == 0.0;
  c->batch_scale[0] = column_scale > 0.0 ? column_scale : 1.0;
  c->batch_scale[1] = block_source_scale > 0.0 ? block_source_scale : 1.0;
  c->user_scales = !reset;
  return 0;
}

}  // extern "C"

struct tgrav_batch {
  tgrav_ctx* c = nullptr;
  int mode = 0, nprob = 0, n_items = 0, nred = 0, grid = 0;
  int n_gitems = 0, ggrid = 0;   // column work list (block targets, explicit targets in C order)
  int waves = 1;                 // waves per column workgroup (the context's, at creation)
  int n_sitems = 0, sgrid = 0;   // block-source work list (k_eval_gridsrc's column walk)
  int ncand = 0;                 // problems whose column structure is tested at every run
  bool skip_generic = false;     // every generic item belongs to a candidate (skipped when all pass)
  const Variant* v = nullptr;
  int64_t total_src = 0, red_total = 0, cand_total = 0, total_sblk = 0;
  int64_t part_bytes = 0;          // chunk partial sums written per run (bytes)
  double inter = 0.0;
  DevBuf probs, items, gitems, sitems, sgoff, srcs, sprefix, s4, se2, sax, be2, bms, stats, part, counter, ids,
      rprefix, cand, cprefix, colflag;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void batch_free(tgrav_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  DevBuf* bufs[] = {&b->probs, &b->items, &b->gitems, &b->sitems, &b->sgoff, &b->srcs, &b->sprefix, &b->s4,
                    &b->se2, &b->sax, &b->be2, &b->bms, &b->stats, &b->part, &b->counter, &b->ids,
                    &b->rprefix, &b->cand, &b->cprefix, &b->colflag};
  for (DevBuf* d : bufs) d->release();
  delete b;
}

// Block-source work items: tiles of K x 64 targets (gs_k), the eager kernel's shape.

// The persistent column kernel for W waves per workgroup.
static const void* batch_grid_kernel(bool mixed, int W) {
  if (W == 4) return mixed ? (const void*)tg::k_batch_eval_grid<true, 4> : (const void*)tg::k_batch_eval_grid<false, 4>;
  if (W == 2) return mixed ? (const void*)tg::k_batch_eval_grid<true, 2> : (const void*)tg::k_batch_eval_grid<false, 2>;
  return mixed ? (const void*)tg::k_batch_eval_grid<true, 1> : (const void*)tg::k_batch_eval_grid<false, 1>;
}

template <bool MIXED, int W>
static void batch_gridsrc_launch(tgrav_batch* b, hipStream_t s) {
  tg::k_batch_eval_gridsrc<gs_k<MIXED>(), W, MIXED><<<b->sgrid, 64 * W, 0, s>>>(
      b->probs.as<tg::BProb>(), b->sitems.as<tg::BItem>(), b->sgoff.as<int>(), b->counter.as<unsigned int>() + 4,
      b->bms.as<double>(), b->sax.as<double>(), b->be2.as<double>(), b->stats.as<unsigned long long>(),
      b->part.as<double>());
}

static const void* batch_gridsrc_kernel(bool mixed, int W) {
  if (W == 4) return mixed ? (const void*)tg::k_batch_eval_gridsrc<gs_k<true>(), 4, true> : (const void*)tg::k_batch_eval_gridsrc<gs_k<false>(), 4, false>;
  if (W == 2) return mixed ? (const void*)tg::k_batch_eval_gridsrc<gs_k<true>(), 2, true> : (const void*)tg::k_batch_eval_gridsrc<gs_k<false>(), 2, false>;
  return mixed ? (const void*)tg::k_batch_eval_gridsrc<gs_k<true>(), 1, true> : (const void*)tg::k_batch_eval_gridsrc<gs_k<false>(), 1, false>;
}

template <bool MIXED, int W>
static void batch_grid_launch(tgrav_batch* b, hipStream_t s) {
  tg::k_batch_eval_grid<MIXED, W><<<b->ggrid, 64 * W, 0, s>>>(
      b->probs.as<tg::BProb>(), b->gitems.as<tg::BItem>(), b->n_gitems, b->counter.as<unsigned int>() + 1,
      b->s4.as<double4>(), b->se2.as<double>(), b->stats.as<unsigned long long>(), b->colflag.as<int>(),
      b->part.as<double>());
}

template <class T>
static int upload(DevBuf& d, const std::vector<T>& v) {
  int rc = d.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (rc) return rc;
  if (!v.empty()) HIPCHECK(hipMemcpy(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Enqueues reset -> pack -> column check -> persistent evals -> chunk reduction on `s`.
static int batch_enqueue(tgrav_batch* b, hipStream_t s, bool timed) {
  tg::k_batch_reset<<<grid1(std::max(b->nprob, 1), 256), 256, 0, s>>>(
      b->nprob, b->counter.as<unsigned int>(), b->stats.as<unsigned long long>(), b->colflag.as<int>(),
      b->skip_generic ? 1u : 0u);
  LAUNCHCHECK("k_batch_reset");
  if (b->total_src > 0) {
    tg::k_batch_pack<<<grid1(b->total_src, 256), 256, 0, s>>>(
        b->nprob, b->srcs.as<tg::BSrc>(), b->sprefix.as<int64_t>(), b->s4.as<double4>(),
        b->se2.as<double>(), b->sax.as<double>(), b->be2.as<double>(), b->bms.as<double>(),
        b->stats.as<unsigned long long>());
    LAUNCHCHECK("k_batch_pack");
  }
  if (b->ncand > 0 && b->cand_total > 0) {
    tg::k_batch_check_cols<<<grid1(b->cand_total, 256), 256, 0, s>>>(
        b->ncand, b->cand.as<int>(), b->cprefix.as<int64_t>(), b->probs.as<tg::BProb>(), b->colflag.as<int>(),
        b->counter.as<unsigned int>());
    LAUNCHCHECK("k_batch_check_cols");
  }
  if (b->n_items > 0) {
    auto go = [&]() -> int {
      b->v->bfn(b->grid, s, b->probs.as<tg::BProb>(), b->items.as<tg::BItem>(), b->n_items,
                b->counter.as<unsigned int>(), b->s4.as<double4>(), b->se2.as<double>(),
                b->stats.as<unsigned long long>(), b->colflag.as<int>(), b->part.as<double>());
      LAUNCHCHECK("k_batch_eval");
      return 0;
    };
    int rc = timed ? timed_launch_on(b->c, s, b->n_items, b->total_src, go) : go();
    if (rc) return rc;
  }
  if (b->n_gitems > 0) {
    auto go = [&]() -> int {
      const bool mixed = b->mode == TGRAV_MIXED;
      switch (b->waves) {
        case 4: mixed ? batch_grid_launch<true, 4>(b, s) : batch_grid_launch<false, 4>(b, s); break;
        case 2: mixed ? batch_grid_launch<true, 2>(b, s) : batch_grid_launch<false, 2>(b, s); break;
        default: mixed ? batch_grid_launch<true, 1>(b, s) : batch_grid_launch<false, 1>(b, s);
      }
      LAUNCHCHECK("k_batch_eval_grid");
      return 0;
    };
    int rc = timed ? timed_launch_on(b->c, s, b->n_gitems, b->total_src, go) : go();
    if (rc) return rc;
  }
  if (b->n_sitems > 0) {
    auto go = [&]() -> int {
      const bool mixed = b->mode == TGRAV_MIXED;
      switch (b->waves) {
        case 4: mixed ? batch_gridsrc_launch<true, 4>(b, s) : batch_gridsrc_launch<false, 4>(b, s); break;
        case 2: mixed ? batch_gridsrc_launch<true, 2>(b, s) : batch_gridsrc_launch<false, 2>(b, s); break;
        default: mixed ? batch_gridsrc_launch<true, 1>(b, s) : batch_gridsrc_launch<false, 1>(b, s);
      }
      LAUNCHCHECK("k_batch_eval_gridsrc");
      return 0;
    };
    int rc = timed ? timed_launch_on(b->c, s, b->n_sitems, b->total_src, go) : go();
    if (rc) return rc;
  }
  if (b->nred > 0 && b->red_total > 0) {
    tg::k_batch_reduce<<<grid1(b->red_total, 256), 256, 0, s>>>(
        b->nred, b->ids.as<int>(), b->rprefix.as<int64_t>(), b->probs.as<tg::BProb>(),
        b->part.as<double>());
    LAUNCHCHECK("k_batch_reduce");
  }
  return 0;
}

extern "C" {

int tgrav_batch_create(tgrav_ctx* c, int32_t n, const tgrav_problem* pr, int mode, double dt,
                       tgrav_batch** out) {
  if (out) *out = nullptr;
  if (!c || !out) return fail(TGRAV_EINVAL, "context and out are required");
  int rc = check_mode(mode);
  if (rc) return rc;
  if (c->soft != TGRAV_SOFT_PLUMMER || c->periodic)
    return fail(TGRAV_EINVAL, "batch plans use the Plummer kernel (tgrav_set_softening(ctx, 0, 0, 0))");
  if (n < 0 || (n > 0 && !pr)) return fail(TGRAV_EINVAL, "bad problem list");
  if (!isfinite(dt)) return fail(TGRAV_EINVAL, "dt must be finite");
  const bool kick = dt != 0.0;
  for (int p = 0; p < n; ++p) {
    const tgrav_problem& q = pr[p];
    if (q.n_sources < 0 || q.n_targets < 0)
      return fail(TGRAV_EINVAL, "problem %d: negative size", p);
    if (q.n_source_blocks < 0) return fail(TGRAV_EINVAL, "problem %d: negative source block count", p);
    if (q.n_source_blocks > 0) {
      if (q.n_sources != 512 * q.n_source_blocks)
        return fail(TGRAV_EINVAL, "problem %d: n_sources must be 512 * n_source_blocks", p);
      if (!q.sbx || !q.sby || !q.sbz || !q.sbd || !q.sm)
        return fail(TGRAV_EINVAL, "problem %d: source block corners, widths and cell masses are required", p);
      if (q.n_source_blocks > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "problem %d: too many source blocks", p);
    } else if (q.n_sources > 0 && (!q.sx || !q.sy || !q.sz || !q.sm)) {
      return fail(TGRAV_EINVAL, "problem %d: source x, y, z and m are required", p);
    }
    if (q.n_target_blocks < 0) return fail(TGRAV_EINVAL, "problem %d: negative block count", p);
    if (q.n_target_blocks > 0) {
      if (q.n_targets != 512 * q.n_target_blocks)
        return fail(TGRAV_EINVAL, "problem %d: n_targets must be 512 * n_target_blocks", p);
      if (!q.tbx || !q.tby || !q.tbz || !q.tbd)
        return fail(TGRAV_EINVAL, "problem %d: block corners and widths are required", p);
      if (q.n_target_blocks > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "problem %d: too many blocks", p);
    } else if (q.n_targets > 0 && (!q.tx || !q.ty || !q.tz)) {
      return fail(TGRAV_EINVAL, "problem %d: target x, y and z are required", p);
    }
    if (kick && q.n_targets > 0 && (!q.vx || !q.vy || !q.vz))
      return fail(TGRAV_EINVAL, "problem %d: velocities are required when dt != 0", p);
  }
  if ((rc = set_device(c))) return rc;
  tgrav_batch* b = new tgrav_batch();
  b->c = c;
  b->mode = mode;
  b->waves = c->col_waves;
  b->nprob = n;
  b->v = variant_of(mode, c->variant[mode]);
  double total = 0.0;
  for (int p = 0; p < n; ++p) total += (double)pr[p].n_sources * (double)pr[p].n_targets;
  b->inter = total;
  double item = c->batch_item;
  int wg_per_cu = c->batch_wg_per_cu;
  double col_scale = c->batch_scale[0], gs_scale = c->batch_scale[1];
  bool tuned_item = c->user_batch && item > 0;
  if (!c->tune.empty() && n > 0) {   // the autotuner's shape for batches like this one
    std::vector<double> vs(n), vt(n);
    for (int p = 0; p < n; ++p) { vs[p] = (double)pr[p].n_sources; vt[p] = (double)pr[p].n_targets; }
    auto median = [](std::vector<double> v) {
      std::sort(v.begin(), v.end());
      const size_t m = v.size() / 2;
      return v.size() % 2 ? v[m] : 0.5 * (v[m - 1] + v[m]);
    };
    char key[256];
    // plans tuned for more than one wave per column workgroup carry it in the key
    char bk[32];
    if (c->col_waves > 1) snprintf(bk, sizeof bk, "batch/w%d", c->col_waves);
    else snprintf(bk, sizeof bk, "batch");
    snprintf(key, sizeof key, "%s|%s|%s|p%d|i%d|s%d|t%d", c->devkey.c_str(), bk, mode_name(mode), n,
             size_bucket(total), size_bucket(median(vs)), size_bucket(median(vt)));
    const TuneEntry* t = tune_lookup(c, key);
    if (t) {
      if (!c->user_cfg[mode] && t->variant >= 0 && variant_of(mode, t->variant)) b->v = variant_of(mode, t->variant);
      if (!c->user_batch) {
        if (t->item >= 0) item = t->item;
        if (t->wg_per_cu >= 0) wg_per_cu = t->wg_per_cu;
        tuned_item = item > 0;
      }
      if (!c->user_scales && t->col_scale > 0) col_scale = t->col_scale;
      if (!c->user_scales && t->gs_scale > 0) gs_scale = t->gs_scale;
    }
  }
  // Resident workgroups of each persistent kernel: the occupancy limit, or for the generic kernel
  // the tuned / set workgroups per CU (the column and block-source kernels fill their CUs).
  auto occupancy = [&](const void* k, int threads, int limit) {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k, threads, 0) != hipSuccess || o < 1) o = 1;
    const int per_cu = limit > 0 ? std::min(limit, o) : o;
    return per_cu * c->num_cu;
  };
  const int slots = occupancy(b->v->bkernel(), b->v->nt, wg_per_cu);
  const int CW = c->col_waves;
  const int gslots = occupancy(batch_grid_kernel(mode == TGRAV_MIXED, CW), 64 * CW, 0);
  const int sslots = occupancy(batch_gridsrc_kernel(mode == TGRAV_MIXED, CW), 64 * CW, 0);
  const int64_t GT = (int64_t)gs_k_mode(mode) * 64;   // targets per block-source tile
  // Form of each problem's work items.
  std::vector<int> form(n, tg::FORM_GENERIC);
  double tot_list[3] = {0.0, 0.0, 0.0};   // interactions on the generic, column, block-source lists
  for (int p = 0; p < n; ++p) {
    const tgrav_problem& q = pr[p];
    if (q.n_source_blocks > 0 && c->block_kernel) form[p] = tg::FORM_GRIDSRC;
    else if (q.n_target_blocks > 0) form[p] = tg::FORM_COLS;
    else if (c->block_kernel && q.n_targets >= 512 && q.n_targets % 512 == 0) form[p] = tg::FORM_CAND;
    const double w = (double)q.n_sources * (double)q.n_targets;
    tot_list[form[p] == tg::FORM_GRIDSRC ? 2 : form[p] == tg::FORM_GENERIC ? 0 : 1] += w;
  }
  const int64_t T = (int64_t)b->v->k * b->v->nt, NT = b->v->nt;
  const double C = item > 0 ? item : std::max((double)T * NT * 4.0, total / (c->num_cu * 8.0));
  // Without a tuned or set item size, each list is cut into about 8 rounds of its resident
  // workgroups at least (equal-cost items balance the last round), down to a floor of 4 source
  // tiles (4 blocks) per item.
  auto list_cost = [&](int li, double floor_cost) {
    if (tuned_item) return C;
    const int sl = li == 0 ? slots : li == 1 ? gslots : sslots;
    return std::min(C, std::max(floor_cost, tot_list[li] / (8.0 * sl)));
  };
  const double C_gen = list_cost(0, (double)T * NT * 4.0), C_col = col_scale * list_cost(1, 512.0 * 64 * 4 * CW),
               C_gs = gs_scale * list_cost(2, (double)GT * 512.0 * 4 * CW);
  std::vector<tg::BProb> probs(n);
  std::vector<tg::BSrc> srcs(n);
  std::vector<int64_t> sprefix(n + 1, 0);
  std::vector<tg::BItem> items, gitems, sitems;
  std::vector<double> cost, gcost, scost;
  std::vector<int> ids, cand;
  std::vector<int64_t> rprefix(1, 0), cprefix(1, 0);
  int64_t part_total = 0, sblk_total = 0;
  for (int p = 0; p < n; ++p) {
    const tgrav_problem& q = pr[p];
    tg::BProb& P = probs[p];
    memset(&P, 0, sizeof P);
    P.form = form[p];
    if (q.n_target_blocks > 0) {
      P.T.blocks = 1; P.T.nb = 8;
      P.T.ox = q.tbx; P.T.oy = q.tby; P.T.oz = q.tbz; P.T.bdx = q.tbd; P.T.beps = q.tbe;
    } else {
      P.T.x = q.tx; P.T.y = q.ty; P.T.z = q.tz; P.T.eps = q.te;
    }
    P.T.n = q.n_targets;
    P.O = tg::Out{q.ax, q.ay, q.az, q.phi, kick ? q.vx : nullptr, kick ? q.vy : nullptr,
                  kick ? q.vz : nullptr, dt, nullptr, 0, c->G, q.phi == nullptr};
    P.src_off = sprefix[p];
    P.ns = q.n_sources;
    if (q.n_source_blocks > 0) {
      srcs[p] = tg::BSrc{nullptr, nullptr, nullptr, q.sm, q.sbe, q.sbx, q.sby, q.sbz, q.sbd, sblk_total};
      P.sblk_off = sblk_total;
      P.nsblk = q.n_source_blocks;
      sblk_total += q.n_source_blocks;
    } else {
      srcs[p] = tg::BSrc{q.sx, q.sy, q.sz, q.sm, q.se, nullptr, nullptr, nullptr, nullptr, 0};
    }
    sprefix[p + 1] = sprefix[p] + q.n_sources;
    if (P.form == tg::FORM_CAND) {     // column test at every run: one lane per target
      cand.push_back(p);
      cprefix.push_back(cprefix.back() + q.n_targets);
    }
    if (q.n_targets == 0) { P.nchunk = 1; continue; }
    if (q.n_sources == 0) {                    // zero field: written by the reduction
      P.nchunk = 0;
      ids.push_back(p);
      rprefix.push_back(rprefix.back() + q.n_targets);
      continue;
    }
    const bool gs = P.form == tg::FORM_GRIDSRC;
    const bool col = P.form == tg::FORM_COLS || P.form == tg::FORM_CAND;
    // Chunking: GRIDSRC in source blocks; otherwise in sources, a multiple of the LDS tile of every
    // kernel that may take the problem's items (64 columns, NT generic; NT is a multiple of 64) --
    // NT for all, so that targets in the blocks' C order and the same targets as blocks get the
    // same chunks, hence the same bits.
    const int64_t TT = gs ? GT : col ? 512 : T;
    const int64_t SG = gs ? 1 : NT;
    const int64_t units = gs ? q.n_source_blocks : q.n_sources;
    const double per_tile = (double)std::min<int64_t>(TT, q.n_targets) * q.n_sources;
    const double Cp = gs ? C_gs : col ? C_col : C_gen;
    int64_t nch = per_tile > Cp ? (int64_t)ceil(per_tile / Cp) : 1;
    // A column chunk's sources (40 B each with eps^2) stay within half an XCD's L2, so that the
    // workgroups in flight, which walk the same chunk out of step, share it there: kColChunk
    // sources (whole-block items over 100,000 sources read 12x the bytes from beyond the L2).
    const int64_t kColChunk = ((int64_t)(2 << 20) / 40) / NT * NT;
    if (col) nch = std::max<int64_t>(nch, (units + kColChunk - 1) / kColChunk);
    int64_t ch = (units + nch - 1) / nch;
    ch = (ch + SG - 1) / SG * SG;
    nch = (units + ch - 1) / ch;
    if (nch > 0x7fffffffLL) { batch_free(b); return fail(TGRAV_ELIMIT, "problem %d: too many chunks", p); }
    P.chunk = ch;
    P.nchunk = (int)nch;
    if (nch > 1) {
      P.part_off = part_total;
      part_total += nch * 4 * q.n_targets;
      ids.push_back(p);
      rprefix.push_back(rprefix.back() + q.n_targets);
    }
    const double per_unit = gs ? 512.0 : 1.0;
    // Items in tile-major order (a tile's chunks next to each other: the generic list), or
    // chunk-major (all tiles of one source chunk next to each other, so that the workgroups
    // running at the same time share the chunk's sources in L2: the column and block-source
    // lists).  The cost sort below is stable, so equal-cost items keep this order.
    auto add = [&](std::vector<tg::BItem>& it, std::vector<double>& cs, int64_t tt, bool chunk_major) -> bool {
      const int64_t tiles = (q.n_targets + tt - 1) / tt;
      if (tiles > 0x7fffffffLL) return false;
      for (int64_t a = 0; a < (chunk_major ? nch : tiles); ++a)
        for (int64_t z = 0; z < (chunk_major ? tiles : nch); ++z) {
          const int64_t t = chunk_major ? z : a, k = chunk_major ? a : z;
          const double nt_t = (double)std::min(tt, q.n_targets - t * tt);
          const double ns_k = per_unit * (double)std::min(ch, units - k * ch);
          it.push_back(tg::BItem{p, (int)t, (int)k, 0});
          cs.push_back(nt_t * ns_k);
        }
      return true;
    };
    bool ok = true;
    if (gs) ok = add(sitems, scost, GT, true);
    if (col) ok = ok && add(gitems, gcost, 512, true);
    if (P.form == tg::FORM_GENERIC || P.form == tg::FORM_CAND) ok = ok && add(items, cost, T, false);
    if (!ok) { batch_free(b); return fail(TGRAV_ELIMIT, "problem %d too large", p); }
  }
  if (items.size() > 0x7fffffffULL || gitems.size() > 0x7fffffffULL || sitems.size() > 0x7fffffffULL) {
    batch_free(b);
    return fail(TGRAV_ELIMIT, "too many work items");
  }
  auto by_cost = [](const std::vector<tg::BItem>& it, const std::vector<double>& cs) {
    std::vector<size_t> order(it.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t z) { return cs[a] > cs[z]; });
    std::vector<tg::BItem> out(it.size());
    for (size_t i = 0; i < order.size(); ++i) out[i] = it[order[i]];
    return out;
  };
  const std::vector<tg::BItem> sorted = by_cost(items, cost), gsorted = by_cost(gitems, gcost),
                               ssorted_all = by_cost(sitems, scost);
  // Block-source items in 8 groups by chunk mod 8 (each group keeps the cost order), so that the
  // workgroups on one XCD share their chunks' masses in its L2 (k_batch_eval_gridsrc).
  std::vector<tg::BItem> ssorted;
  std::vector<int> sgoff(9, 0);
  ssorted.reserve(ssorted_all.size());
  for (int q = 0; q < 8; ++q) {
    sgoff[q] = (int)ssorted.size();
    for (const tg::BItem& w : ssorted_all)
      if ((w.chunk & 7) == q) ssorted.push_back(w);
  }
  sgoff[8] = (int)ssorted.size();
  b->n_items = (int)sorted.size();
  b->n_gitems = (int)gsorted.size();
  b->n_sitems = (int)ssorted.size();
  b->nred = (int)ids.size();
  b->red_total = rprefix.back();
  b->total_src = sprefix[n];
  b->total_sblk = sblk_total;
  b->part_bytes = part_total * (int64_t)sizeof(double);
  b->ncand = (int)cand.size();
  b->skip_generic = !sorted.empty() && std::all_of(sorted.begin(), sorted.end(), [&](const tg::BItem& w) {
    return probs[w.prob].form == tg::FORM_CAND;
  });
  b->cand_total = cprefix.back();
  b->grid = std::max(1, std::min(b->n_items, slots));
  b->ggrid = std::max(1, std::min(b->n_gitems, gslots));
  b->sgrid = std::max(1, std::min(b->n_sitems, sslots));

  if ((rc = upload(b->probs, probs)) || (rc = upload(b->items, sorted)) ||
      (rc = upload(b->gitems, gsorted)) || (rc = upload(b->sitems, ssorted)) || (rc = upload(b->sgoff, sgoff)) ||
      (rc = upload(b->srcs, srcs)) || (rc = upload(b->sprefix, sprefix)) ||
      (rc = upload(b->ids, ids)) || (rc = upload(b->rprefix, rprefix)) ||
      (rc = upload(b->cand, cand)) || (rc = upload(b->cprefix, cprefix)) ||
      (rc = b->s4.ensure(std::max<int64_t>(b->total_src, 1) * sizeof(double4))) ||
      (rc = b->se2.ensure(std::max<int64_t>(b->total_src, 1) * sizeof(double))) ||
      (rc = b->sax.ensure(std::max<int64_t>(sblk_total, 1) * 24 * sizeof(double))) ||
      (rc = b->be2.ensure(std::max<int64_t>(sblk_total, 1) * sizeof(double))) ||
      (rc = b->bms.ensure(std::max<int64_t>(sblk_total, 1) * 512 * sizeof(double))) ||
      (rc = b->stats.ensure((size_t)std::max(n, 1) * tg::ST_WORDS * sizeof(unsigned long long))) ||
      (rc = b->colflag.ensure((size_t)std::max(n, 1) * sizeof(int))) ||
      (rc = b->part.ensure(std::max<int64_t>(part_total, 1) * sizeof(double))) ||
      (rc = b->counter.ensure(12 * sizeof(unsigned int)))) {
    batch_free(b);
    return rc;
  }
  if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) {
    batch_free(b);
    return fail(TGRAV_EHIP, "cannot create the capture stream");
  }
  *out = b;
  return 0;
}

int tgrav_batch_run(tgrav_batch* b, int use_graph) {
  if (!b) return fail(TGRAV_EINVAL, "batch is NULL");
  tgrav_ctx* c = b->c;
  int rc = set_device(c);
  if (rc) return rc;
  if (!use_graph) return batch_enqueue(b, c->stream, true);
  if (!b->exec) {
    HIPCHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(b, b->cap, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(TGRAV_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      b->exec = nullptr;
      return fail(TGRAV_EHIP, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    b->graph = g;
  }
  HIPCHECK(hipGraphLaunch(b->exec, c->stream));
  return 0;
}

int tgrav_batch_destroy(tgrav_batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->c->device);
  (void)hipStreamSynchronize(b->c->stream);
  batch_free(b);
  return 0;
}

int tgrav_batch_partials(tgrav_batch* b, int64_t* bytes) {
  if (!b || !bytes) return fail(TGRAV_EINVAL, "batch or output is NULL");
  *bytes = b->part_bytes;
  return 0;
}

int tgrav_batch_info(tgrav_batch* b, int64_t* n_items, double* interactions) {
  if (!b) return fail(TGRAV_EINVAL, "batch is NULL");
  if (n_items) *n_items = (int64_t)b->n_items + b->n_gitems;
  if (interactions) *interactions = b->inter;
  return 0;
}

}  // extern "C"
