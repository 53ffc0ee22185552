// runtime.hip -- context, errors, device buffers and kernel timing.
// Replaces the reference's OpenCL set-up (src/clrun.c:51-514 cl_init_device,
// src/clrun.c:1268-1490 cl_release_device): one device, one in-order stream.
#include <stdarg.h>
#include <string.h>

#include <mutex>

#include "runtime.h"

namespace zr {

static Ctx g_ctx;
static char g_err[1024] = "";

Ctx &ctx() { return g_ctx; }

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  if (g_ctx.cfg.verbose) fprintf(stderr, "zeta: %s\n", g_err);
  return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  (void)hipGetLastError();  // a failed call leaves the error behind for the next launch check: clear it
  return fail(ZETA_EHIP, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
}

int require_init() {
  if (!g_ctx.up) return fail(ZETA_ESTATE, "library not initialised (call zeta_init)");
  return ZETA_OK;
}

int require_up() {
  if (!g_ctx.up) return fail(ZETA_ESTATE, "library not initialised (call zeta_init)");
  if (g_ctx.serving) return fail(ZETA_EBUSY, "the device is serving (call zeta_serve_stop first)");
  return ZETA_OK;
}

int ensure(DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return ZETA_OK;
  if (b.p) {
    // work queued on the library stream may still read the old buffer
    if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return hip_fail(e, "hipFree", __FILE__, __LINE__);
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(ZETA_ENOMEM, "hipMalloc of %zu bytes failed (%s)", want, hipGetErrorString(e));
  }
  b.bytes = want;
  return ZETA_OK;
}

void release(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

static const char *kname(int id) {
  static const char *names[K_COUNT_IDS] = {"perft_roots", "perft_count", "perft_scan", "perft_expand",
                                           "perft_apply", "perft_leaf",  "playout",    "search",
                                           "mlp_fwd",     "mlp_bwd",     "mlp_adam",   "mlp_score",
                                           "copy"};
  return names[id];
}

static hipEvent_t get_event() {
  if (!g_ctx.event_pool.empty()) {
    hipEvent_t e = g_ctx.event_pool.back();
    g_ctx.event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void time_begin(int id, uint64_t items, hipEvent_t *a, hipEvent_t *b) {
  *a = *b = nullptr;
  if (!g_ctx.timing) return;
  *a = get_event();
  *b = get_event();
  if (*a) (void)hipEventRecord(*a, g_ctx.stream);
  (void)id;
  (void)items;
}

void time_end(hipEvent_t a, hipEvent_t b, int id, uint64_t items) {
  if (!g_ctx.timing || !a || !b) return;
  (void)hipEventRecord(b, g_ctx.stream);
  g_ctx.pending.push_back({id, a, b, items});
}

void time_collect() {
  for (auto &p : g_ctx.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      zeta_kstat &k = g_ctx.kstat[p.id];
      k.launches += 1;
      k.ms += ms;
      k.items += p.items;
    }
    g_ctx.event_pool.push_back(p.a);
    g_ctx.event_pool.push_back(p.b);
  }
  g_ctx.pending.clear();
}

}  // namespace zr

using namespace zr;

// per-device set-up of the perft subsystem (perft.hip): the address of the
// allocation-free entry point's scratch array on the current device
int perft_setup();

extern "C" int zeta_init(const zeta_config *cfg) {
  Ctx &c = ctx();
  if (c.up) return ZETA_OK;
  int rc = ZETA_OK;
  zeta_config d{};
  if (cfg) d = *cfg;
  if (d.frontier_budget == 0) d.frontier_budget = 1ULL << 28;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(ZETA_ENODEV, "no HIP device available (%s)", e == hipSuccess ? "count 0" : hipGetErrorString(e));
  }
  if (d.device < 0 || d.device >= n) return fail(ZETA_EINVAL, "device %d out of range (%d devices)", d.device, n);
  e = hipSetDevice(d.device);
  if (e != hipSuccess) return fail(ZETA_ENODEV, "hipSetDevice(%d): %s", d.device, hipGetErrorString(e));
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, d.device);
  if (e != hipSuccess) return fail(ZETA_ENODEV, "hipGetDeviceProperties: %s", hipGetErrorString(e));
  c.arch = prop.gcnArchName;
  if (c.arch.compare(0, 6, "gfx950") != 0)
    return fail(ZETA_ENODEV, "device %d is %s; this library is built for gfx950 only", d.device, prop.gcnArchName);
  c.cu_count = prop.multiProcessorCount;
  e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(ZETA_ENODEV, "hipStreamCreate: %s", hipGetErrorString(e));
  void *ab = nullptr;
  e = hipHostMalloc(&ab, 64, hipHostMallocCoherent | hipHostMallocMapped);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    return fail(ZETA_ENOMEM, "hipHostMalloc of the abort word: %s", hipGetErrorString(e));
  }
  memset(ab, 0, 64);
  void *abd = nullptr;
  e = hipHostGetDevicePointer(&abd, ab, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(ab);
    (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    return fail(ZETA_ENODEV, "hipHostGetDevicePointer of the abort word: %s", hipGetErrorString(e));
  }
  c.abort_host = (uint32_t *)ab;
  c.abort_dev = (const uint32_t *)abd;
  c.perft_total_host = (uint64_t *)((char *)ab + 8);
  e = hipEventCreateWithFlags(&c.ev_total, hipEventDisableTiming);
  if (e == hipSuccess && (rc = perft_setup()) != ZETA_OK) e = hipErrorUnknown;
  if (e != hipSuccess) {
    if (c.ev_total) (void)hipEventDestroy(c.ev_total);
    c.ev_total = nullptr;
    (void)hipHostFree(ab);
    c.abort_host = nullptr;
    c.abort_dev = nullptr;
    c.perft_total_host = nullptr;
    (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    return rc ? rc : fail(ZETA_ENODEV, "hipEventCreate: %s", hipGetErrorString(e));
  }
  c.cfg = d;
  c.device = d.device;
  memset(c.kstat, 0, sizeof(c.kstat));
  c.up = true;
  return ZETA_OK;
}

// release hooks of the subsystems (defined in their own files)
void perft_release();
void search_release();
void mlp_release();

extern "C" int zeta_shutdown(void) {
  Ctx &c = ctx();
  if (!c.up) return ZETA_OK;
  if (c.serving) (void)zeta_serve_stop(nullptr);
  (void)hipStreamSynchronize(c.stream);
  time_collect();
  perft_release();
  search_release();
  mlp_release();
  for (auto e : c.event_pool) (void)hipEventDestroy(e);
  c.event_pool.clear();
  if (c.ev_total) (void)hipEventDestroy(c.ev_total);
  c.ev_total = nullptr;
  if (c.abort_host) (void)hipHostFree(c.abort_host);
  c.abort_host = nullptr;
  c.abort_dev = nullptr;
  c.perft_total_host = nullptr;
  (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
  c.up = false;
  return ZETA_OK;
}

extern "C" int zeta_last_error(char *buf, size_t n) {
  if (!buf || n == 0) return ZETA_EINVAL;
  snprintf(buf, n, "%s", zr::g_err);
  return ZETA_OK;
}

extern "C" int zeta_device_info(char *buf, size_t n) {
  int r = require_init();
  if (r) return r;
  if (!buf || n == 0) return ZETA_EINVAL;
  snprintf(buf, n, "%s cu=%d", ctx().arch.c_str(), ctx().cu_count);
  return ZETA_OK;
}

extern "C" int zeta_set_timing(int on) {
  int r = require_init();
  if (r) return r;
  ctx().timing = on != 0;
  return ZETA_OK;
}

extern "C" int zeta_stats_get(zeta_stats *st) {
  int r = require_up();
  if (r) return r;
  if (!st) return ZETA_EINVAL;
  (void)hipStreamSynchronize(ctx().stream);
  time_collect();
  memset(st, 0, sizeof(*st));
  for (int i = 0; i < K_COUNT_IDS && i < ZETA_MAX_KSTATS; i++) {
    st->k[i] = ctx().kstat[i];
    snprintf(st->k[i].name, sizeof(st->k[i].name), "%s", kname(i));
  }
  st->nk = K_COUNT_IDS < ZETA_MAX_KSTATS ? K_COUNT_IDS : ZETA_MAX_KSTATS;
  return ZETA_OK;
}

extern "C" int zeta_stats_reset(void) {
  int r = require_up();
  if (r) return r;
  (void)hipStreamSynchronize(ctx().stream);
  time_collect();
  memset(ctx().kstat, 0, sizeof(ctx().kstat));
  return ZETA_OK;
}

extern "C" int zeta_dev_alloc(void **dptr, size_t bytes) {
  int r = require_up();
  if (r) return r;
  if (!dptr) return fail(ZETA_EINVAL, "zeta_dev_alloc: null out pointer");
  *dptr = nullptr;
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return fail(ZETA_ENOMEM, "hipMalloc of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
  }
  return ZETA_OK;
}

extern "C" int zeta_dev_free(void *dptr) {
  int r = require_up();
  if (r) return r;
  if (!dptr) return ZETA_OK;
  ZCHECK(hipStreamSynchronize(ctx().stream));
  ZCHECK(hipFree(dptr));
  return ZETA_OK;
}

extern "C" int zeta_memcpy_htod(void *dst, const void *src, size_t bytes) {
  int r = require_up();
  if (r) return r;
  if (bytes && (!dst || !src)) return fail(ZETA_EINVAL, "zeta_memcpy_htod: null pointer");
  ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx().stream));
  ZCHECK(hipStreamSynchronize(ctx().stream));
  return ZETA_OK;
}

extern "C" int zeta_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  int r = require_up();
  if (r) return r;
  if (bytes && (!dst || !src)) return fail(ZETA_EINVAL, "zeta_memcpy_dtoh: null pointer");
  ZCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().stream));
  ZCHECK(hipStreamSynchronize(ctx().stream));
  return ZETA_OK;
}

// streaming copy, 16 B per lane per iteration (nontemporal: the data is not reused)
typedef unsigned int cu32x4 __attribute__((ext_vector_type(4)));
__global__ void k_copy16(const cu32x4 *__restrict__ src, cu32x4 *__restrict__ dst, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
    const cu32x4 v = __builtin_nontemporal_load(src + i);
    __builtin_nontemporal_store(v, dst + i);
  }
}

extern "C" int zeta_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  int r = require_up();
  if (r) return r;
  if (!bytes) return ZETA_OK;
  if (!dst || !src) return fail(ZETA_EINVAL, "zeta_memcpy_dtod: null pointer");
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return fail(ZETA_EINVAL, "zeta_memcpy_dtod: pointers and size must be multiples of 16 bytes");
  const size_t n16 = bytes / 16;
  size_t blocks = (n16 + 255) / 256;
  const size_t cap = (size_t)ctx().cu_count * 8;
  if (blocks > cap) blocks = cap;
  ZLAUNCH(K_COPY, bytes, k_copy16, (unsigned)blocks, 256, 0, (const cu32x4 *)src, (cu32x4 *)dst, n16);
  ZCHECK(hipStreamSynchronize(ctx().stream));
  time_collect();
  return ZETA_OK;
}

extern "C" int zeta_synchronize(void) {
  int r = require_up();
  if (r) return r;
  ZCHECK(hipStreamSynchronize(ctx().stream));
  time_collect();
  return ZETA_OK;
}
