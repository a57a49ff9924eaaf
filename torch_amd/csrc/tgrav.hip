// tgrav: C ABI of the MI355X direct-sum gravity field (see include/tgrav.h).
//
// Host side: argument checking, host<->device staging for TGRAV_HOST pointers, source
// packing, launch-shape selection (kernel variant + source chunking) and the chunk
// reduction.  Device code lives in kernels.hpp (grid kernels) and batch.hpp
// (persistent batch kernel).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/tgrav.h"
#include "kernels.hpp"

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

template <int K, int NT>
void launch_fp64(dim3 g, dim3 b, hipStream_t s, const tg::Tgt& T, const double4* s4,
                 const double* se2, int64_t ns, int64_t chunk, const double* e2m, const tg::Out& O) {
  tg::k_eval_fp64<K, NT><<<g, b, 0, s>>>(T, s4, se2, ns, chunk, e2m, O);
}
template <int K, int NT, int FOLD>
void launch_mixed(dim3 g, dim3 b, hipStream_t s, const tg::Tgt& T, const double4* s4,
                  const double* se2, int64_t ns, int64_t chunk, const double* e2m, const tg::Out& O) {
  tg::k_eval_mixed<K, NT, FOLD><<<g, b, 0, s>>>(T, s4, se2, ns, chunk, e2m, O);
}

struct Variant {
  int k, nt, ts;
  launch_fn fn;
};

const Variant kFp64Variants[] = {
    {1, 256, 256, launch_fp64<1, 256>}, {2, 256, 256, launch_fp64<2, 256>},
    {4, 256, 256, launch_fp64<4, 256>}, {2, 128, 128, launch_fp64<2, 128>},
    {4, 128, 128, launch_fp64<4, 128>}, {8, 64, 64, launch_fp64<8, 64>},
};
const Variant kMixedVariants[] = {
    {2, 256, 256, launch_mixed<2, 256, 64>}, {4, 256, 256, launch_mixed<4, 256, 64>},
    {8, 256, 256, launch_mixed<8, 256, 64>}, {4, 128, 128, launch_mixed<4, 128, 64>},
    {8, 128, 128, launch_mixed<8, 128, 64>}, {16, 64, 64, launch_mixed<16, 64, 64>},
};
const int kNumFp64 = sizeof(kFp64Variants) / sizeof(kFp64Variants[0]);
const int kNumMixed = sizeof(kMixedVariants) / sizeof(kMixedVariants[0]);
const int kDefaultFp64 = 1;
const int kDefaultMixed = 1;

const Variant* variant_of(int mode, int v) {
  if (mode == TGRAV_PLAIN_FP64) return (v >= 0 && v < kNumFp64) ? &kFp64Variants[v] : nullptr;
  if (mode == TGRAV_MIXED) return (v >= 0 && v < kNumMixed) ? &kMixedVariants[v] : nullptr;
  return nullptr;
}

}  // namespace

struct tgrav_ctx {
  int device = 0;
  int num_cu = 256;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  double G = 4.300917270e-3;
  int validate = 1;
  int variant[2] = {kDefaultFp64, kDefaultMixed};
  int nsplit[2] = {0, 0};
  // packed sources; stat[0] = bits of min eps_s^2 (read by the kernels), stat[1] = NaN flag
  int64_t ns = -1;
  DevBuf s4, se2, stat;
  // staging for host pointers and scratch; flag[0] = NaN flag of the targets
  DevBuf in[8], outb[4], vel[3], part, flag;
  unsigned long long* host_flag = nullptr;   // pinned
};

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

static unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Reads a device flag word written by k_pack_sources / k_check_finite (synchronous).
static int read_flag(tgrav_ctx* c, const void* dflag, bool* bad) {
  HIPCHECK(hipMemcpyAsync(c->host_flag, dflag, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  *bad = *c->host_flag != 0ULL;
  return 0;
}

// NaN/Inf check of device target arrays (synchronous; only with validation on).
static int device_check_finite(tgrav_ctx* c, int64_t n, const double* const* chk, int nchk,
                               bool* bad) {
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
                         int64_t* ntiles, int* nsplit, int64_t* chunk) {
  const int64_t tile = (int64_t)v->k * v->nt;
  *ntiles = (nt + tile - 1) / tile;
  int want = c->nsplit[mode];
  if (want <= 0) {
    const int64_t target_wg = 4LL * c->num_cu;
    if (*ntiles >= target_wg / 2) want = 1;
    else {
      int64_t s = (target_wg + *ntiles - 1) / *ntiles;
      int64_t maxs = (ns + 4LL * v->ts - 1) / (4LL * v->ts);   // >= 4 source tiles a chunk
      if (s > maxs) s = maxs;
      want = (int)(s < 1 ? 1 : s);
    }
  }
  if (want > 65535) want = 65535;
  int64_t ch = (ns + want - 1) / want;
  ch = (ch + v->ts - 1) / v->ts * v->ts;
  if (ch < v->ts) ch = v->ts;
  *chunk = ch;
  *nsplit = (int)((ns + ch - 1) / ch);
  if (*nsplit < 1) *nsplit = 1;
}

// Core launch: targets T (device), outputs O (device), on c->stream.
static int launch_eval(tgrav_ctx* c, int mode, const tg::Tgt& T, const tg::Out& O) {
  const int64_t nt = T.n, ns = c->ns;
  if (nt == 0) return 0;
  // Every pointer a kernel dereferences must be set before anything is launched.
  if (T.blocks ? (!T.ox || !T.oy || !T.oz || !T.bdx) : (!T.x || !T.y || !T.z))
    return fail(TGRAV_EINVAL, "internal: target arrays not bound");
  if (O.vx && (!O.vy || !O.vz)) return fail(TGRAV_EINVAL, "internal: partial velocity set");
  if (ns > 0 && (!c->s4.p || !c->se2.p || !c->stat.p))
    return fail(TGRAV_EINVAL, "internal: source buffers missing");
  if (ns == 0) {   // no sources: zero field, kick adds nothing
    tg::k_fill<<<grid1(nt, 256), 256, 0, c->stream>>>(nt, 0.0, O.ax, O.ay, O.az, O.phi);
    LAUNCHCHECK("k_fill");
    return 0;
  }
  const Variant* v = variant_of(mode, c->variant[mode]);
  int64_t ntiles, chunk;
  int nsplit;
  choose_split(c, mode, v, nt, ns, &ntiles, &nsplit, &chunk);
  if (ntiles > 0x7fffffffLL) return fail(TGRAV_ELIMIT, "too many targets (%lld)", (long long)nt);
  tg::Out o = O;
  if (nsplit > 1) {
    int rc = c->part.ensure((size_t)nsplit * 4 * nt * sizeof(double));
    if (rc) return rc;
    o.part = c->part.as<double>();
    o.pstride = nt;
  } else {
    o.part = nullptr;
  }
  v->fn(dim3((unsigned)ntiles, (unsigned)nsplit), dim3(v->nt), c->stream, T, c->s4.as<double4>(),
        c->se2.as<double>(), ns, chunk, c->stat.as<const double>(), o);
  LAUNCHCHECK(mode == TGRAV_MIXED ? "k_eval_mixed" : "k_eval_fp64");
  if (nsplit > 1) {
    tg::Out fin = O;
    fin.part = nullptr;
    tg::k_reduce_chunks<<<grid1(nt, 256), 256, 0, c->stream>>>(c->part.as<double>(), nt, nsplit,
                                                                nt, fin);
    LAUNCHCHECK("k_reduce_chunks");
  }
  return 0;
}

// ------------------------------------------------------------------- C ABI -------
extern "C" {

int tgrav_version(void) { return TGRAV_VERSION; }

const char* tgrav_last_error(void) { return g_err; }

int tgrav_create(tgrav_ctx** out, int device) {
  if (!out) return fail(TGRAV_EINVAL, "tgrav_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIPCHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(TGRAV_EINVAL, "tgrav_create: device %d out of range (%d devices)", device, ndev);
  tgrav_ctx* c = new tgrav_ctx();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->num_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  }
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&c->host_flag, sizeof(unsigned long long), hipHostMallocDefault);
  if (e != hipSuccess) {
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return fail(TGRAV_EHIP, "tgrav_create: %s", hipGetErrorString(e));
  }
  c->stream = c->own;
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
  c->flag.release();
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
  if (n == 0) {
    c->ns = 0;
    return 0;
  }
  if ((rc = c->s4.ensure((size_t)n * sizeof(double4)))) return rc;
  if ((rc = c->se2.ensure((size_t)n * sizeof(double)))) return rc;
  if ((rc = c->stat.ensure(2 * sizeof(unsigned long long)))) return rc;
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
  if (kind == TGRAV_DEVICE) {
    if (c->validate) {
      bool bad = false;
      if ((rc = read_flag(c, c->stat.as<unsigned long long>() + 1, &bad))) return rc;
      if (bad) return fail(TGRAV_ENONFINITE, "non-finite value in a source array");
    }
  } else {
    HIPCHECK(hipStreamSynchronize(c->stream));   // host arrays may be freed on return
  }
  c->ns = n;
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
    tg::Out O{dout[0], dout[1], dout[2], dout[3], dvel[0], dvel[1], dvel[2], dt, nullptr, 0, c->G};
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
  tg::Out O{outs[0], outs[1], outs[2], outs[3], vel[0], vel[1], vel[2], dt, nullptr, 0, c->G};
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

int tgrav_bidirectional(tgrav_ctx* c, int64_t, const double*, const double*, const double*,
                        const double*, const double*, int64_t, const double*, const double*,
                        const double*, const double*, const double*, double*, double*, double*,
                        double*, double*, double*, double*, double*, int, int) {
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  return fail(TGRAV_ELIMIT, "tgrav_bidirectional: not built yet");
}

int tgrav_batch_create(tgrav_ctx* c, int32_t, const tgrav_problem*, int, double, tgrav_batch** out) {
  if (out) *out = nullptr;
  if (!c) return fail(TGRAV_EINVAL, "context is NULL");
  return fail(TGRAV_ELIMIT, "tgrav_batch_create: not built yet");
}
int tgrav_batch_run(tgrav_batch*, int) { return fail(TGRAV_ELIMIT, "tgrav_batch_run: not built yet"); }
int tgrav_batch_destroy(tgrav_batch*) { return 0; }
int tgrav_batch_info(tgrav_batch*, int64_t*, double*) { return fail(TGRAV_ELIMIT, "not built yet"); }

}  // extern "C"
