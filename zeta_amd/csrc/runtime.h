// runtime.h -- library context, error reporting and per-kernel timing for
// libzetahip.so (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "../../include/zeta_hip.h"

namespace zr {

// kernel ids for the per-kernel HIP-event timers
enum KernelId {
  K_PERFT_ROOTS = 0,
  K_PERFT_COUNT,
  K_PERFT_SCAN,
  K_PERFT_EXPAND,
  K_PERFT_APPLY,
  K_PERFT_LEAF,
  K_PLAYOUT,
  K_SEARCH,
  K_MLP_FWD,
  K_MLP_BWD,
  K_MLP_ADAM,
  K_MLP_SCORE,
  K_COPY,
  K_COUNT_IDS
};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct Ctx {
  bool up = false;
  int device = 0;
  hipStream_t stream = nullptr;
  zeta_config cfg{};
  std::string arch;
  int cu_count = 0;
  bool timing = false;
  bool serving = false;  // a persistent serving grid owns the stream (zeta_serve_start .. stop)
  // 64 bytes of coherent pinned host memory, allocated by zeta_init:
  // bytes 0-3 the search abort word (zeta_search_abort) that the search
  // workers poll at abort_dev, bytes 8-15 the perft level total the host
  // waits for (perft_total_host, after ev_total)
  uint32_t *abort_host = nullptr;
  const uint32_t *abort_dev = nullptr;
  uint64_t *perft_total_host = nullptr;
  hipEvent_t ev_total = nullptr;
  // timing
  struct Pending {
    int id;
    hipEvent_t a, b;
    uint64_t items;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  zeta_kstat kstat[K_COUNT_IDS]{};
};

Ctx &ctx();
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
int require_up();       // initialised and the stream free (ZETA_EBUSY while serving)
int require_init();     // initialised
int ensure(DevBuf &b, size_t bytes);  // grow-only device buffer
void release(DevBuf &b);

// HIP-event timing around a kernel launch on the context stream
void time_begin(int id, uint64_t items, hipEvent_t *a, hipEvent_t *b);
void time_end(hipEvent_t a, hipEvent_t b, int id, uint64_t items);
void time_collect();  // after a stream sync: fold finished events into kstat

}  // namespace zr

#define ZCHECK(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return zr::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

// launch with optional event timing; checks the launch error
#define ZLAUNCH(id, items, kernel, grid, block, lds, ...)                         \
  do {                                                                            \
    hipEvent_t ta_ = nullptr, tb_ = nullptr;                                      \
    zr::time_begin((id), (items), &ta_, &tb_);                                    \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (lds), zr::ctx().stream, __VA_ARGS__); \
    hipError_t le_ = hipGetLastError();                                           \
    if (le_ != hipSuccess) return zr::hip_fail(le_, #kernel, __FILE__, __LINE__);  \
    zr::time_end(ta_, tb_, (id), (items));                                        \
  } while (0)
