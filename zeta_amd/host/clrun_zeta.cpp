// clrun_zeta.cpp -- the reference's six-function device boundary
// (src/clrun.h:26-31, implemented in src/clrun.c over OpenCL) rebuilt over
// libzetahip.so.  The protocol and the process globals are the reference's;
// the work behind cl_run_perft / cl_run_alphabeta is the batched HIP path
// with a batch of one root.
#include <stdio.h>
#include <string.h>

#include "../../include/zeta_clrun.h"
#include "../../include/zeta_hip.h"

namespace {

constexpr int MAXPLY = 64;  // src/types.h:76

struct ShimState {
  bool up = false;
  bool fresh_tt = true;  // the TT is cleared on the first search after cl_init_device
  zeta_pos root{};
  bool have_root = false;
  uint64_t hist[ZETA_HISTORY] = {0};  // game history before the root, newest first
  int32_t hist_len = 0;
  // results of the last run, copied out by cl_read_memory
  bool have_result = false;
  uint64_t nodes = 0;
  bool complete = true;
  uint64_t budget = 0;
  int32_t score = 0;
  uint32_t bestmove = 0;
  uint32_t pv[MAXPLY - 1] = {0};
  uint64_t tt_hits = 0, tt_cuts = 0;
};
ShimState g;

bool report(const char *what, int rc) {
  char msg[512];
  zeta_last_error(msg, sizeof msg);
  fprintf(stderr, "Error (%s failed, %d): %s\n", what, rc, msg);
  return false;
}

int tt_log2_of(uint64_t mb) {
  if (mb == 0) return 0;  // tt1_memory 0 disables the table (src/clrun.c:398-402)
  const uint64_t entries = mb * 1024ULL * 1024ULL / 16ULL;
  int k = 0;
  while ((2ULL << k) <= entries && k < 34) k++;
  return k;
}

int workers_of() {
  const uint64_t w = totalWorkUnits ? totalWorkUnits : 1;
  return w > 1024 ? 1024 : (int)w;  // zeta_search_opts.workers' range
}

}  // namespace

int32_t zeta_split_depth = 0;
int32_t zeta_waves_per_simd = 0;

extern "C" bool cl_init_device(char *kernelname) {
  (void)kernelname;  // one library serves "perft_gpu" and "alphabeta_gpu"
  if (g.up) return true;
  zeta_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.device = opencl_device_id;
  cfg.tt_log2 = tt_log2_of(tt1_memory);
  cfg.workers = workers_of();
  const int rc = zeta_init(&cfg);
  if (rc) return report("cl_init_device", rc);
  g.up = true;
  g.fresh_tt = true;
  g.have_root = false;
  g.have_result = false;
  return true;
}

extern "C" bool cl_write_objects(void) {
  if (!g.up) return report("cl_write_objects", ZETA_ESTATE);
  if (!GLOBAL_BOARD) {
    fprintf(stderr, "Error (cl_write_objects): GLOBAL_BOARD is not allocated\n");
    return false;
  }
  memset(&g.root, 0, sizeof g.root);
  for (int i = 0; i < 7; i++) g.root.board[i] = GLOBAL_BOARD[i];
  // worker 0's copy of the game history: entries 0..PLY-1 precede the root
  // (src/search.c:58-63, src/zeta.cl:1560-1577)
  g.hist_len = 0;
  if (GLOBAL_HASHHISTORY)
    for (int32_t k = 0; k < ZETA_HISTORY && PLY - 1 - k >= 0; k++) g.hist[g.hist_len++] = GLOBAL_HASHHISTORY[PLY - 1 - k];
  g.have_root = true;
  g.have_result = false;
  return true;
}

extern "C" bool cl_run_perft(bool stm, int32_t depth) {
  if (!g.up || !g.have_root) return report("cl_run_perft", ZETA_ESTATE);
  g.root.stm = stm ? 1 : 0;
  g.nodes = 0;
  g.complete = true;
  g.budget = 0;
  g.score = 0;
  g.bestmove = 0;
  memset(g.pv, 0, sizeof g.pv);
  g.tt_hits = g.tt_cuts = 0;
  if (depth >= 1) {
    uint64_t count = 0;
    const int rc = zeta_perft_batch(&g.root, 1, depth, &count);
    if (rc) return report("cl_run_perft", rc);
    g.nodes = count;
  }
  g.have_result = true;
  return true;
}

extern "C" bool cl_run_alphabeta(bool stm, int32_t depth, uint64_t nodes) {
  if (!g.up || !g.have_root) return report("cl_run_alphabeta", ZETA_ESTATE);
  g.root.stm = stm ? 1 : 0;
  g.nodes = 0;
  g.complete = true;
  g.budget = nodes;
  g.score = 0;
  g.bestmove = 0;
  memset(g.pv, 0, sizeof g.pv);
  g.tt_hits = g.tt_cuts = 0;
  if (depth >= 1) {  // the start-up call runs depth 0 (src/zeta.c:695)
    zeta_search_opts o;
    memset(&o, 0, sizeof o);
    o.depth = depth > 40 ? 40 : depth;
    o.workers = workers_of();
    o.tt_log2 = -1;
    o.clear_tt = g.fresh_tt ? 1 : 0;
    o.max_nodes = nodes;
    o.history = g.hist;
    o.history_len = &g.hist_len;
    o.pv = g.pv;
    o.pv_len = MAXPLY - 1;
    o.split_depth = zeta_split_depth;
    o.waves_per_simd = zeta_waves_per_simd;
    zeta_result r;
    memset(&r, 0, sizeof r);
    const int rc = zeta_search_batch(&g.root, 1, &o, &r);
    if (rc) return report("cl_run_alphabeta", rc);
    g.fresh_tt = false;
    g.nodes = r.nodes;
    g.complete = r.status == ZETA_OK;
    g.score = r.score;
    g.bestmove = r.bestmove;
    uint64_t c[8] = {0};
    if (zeta_search_counters(c, 8) == ZETA_OK) {
      g.tt_hits = c[1];
      g.tt_cuts = c[2];
    }
  }
  g.have_result = true;
  return true;
}

extern "C" bool cl_read_memory(void) {
  if (!g.up || !g.have_result) return report("cl_read_memory", ZETA_ESTATE);
  if (!COUNTERS || !PV) {
    fprintf(stderr, "Error (cl_read_memory): COUNTERS or PV is not allocated\n");
    return false;
  }
  const uint64_t w = totalWorkUnits ? totalWorkUnits : 1;
  memset(COUNTERS, 0, sizeof(uint64_t) * w * threadsZ);
  // worker 0 carries its share of the nodes, as in the reference: the host
  // discards a search whose worker 0 reached the per-worker budget
  // (src/search.c:133); an aborted search reports the budget itself
  const uint64_t share = g.nodes / w;
  uint64_t first = share;
  if (g.budget) {
    if (!g.complete) first = g.budget;
    else if (first >= g.budget) first = g.budget - 1;
  }
  COUNTERS[1] = first;
  uint64_t rest = g.nodes > first ? g.nodes - first : 0;
  for (uint64_t i = 1; i < w && rest; i++) {
    const uint64_t take = (i + 1 == w) ? rest : (share < rest ? share : rest);
    COUNTERS[i * 64 + 1] = take;
    rest -= take;
  }
  if (rest) COUNTERS[1] += rest;
  COUNTERS[3] = g.tt_hits;
  COUNTERS[4] = g.tt_cuts;
  COUNTERS[5] = 0;
  memset(PV, 0, sizeof(uint32_t) * MAXPLY);
  PV[0] = (uint32_t)g.score;
  PV[1] = g.bestmove;
  for (int k = 1; k < MAXPLY - 1 && g.pv[k]; k++) PV[k + 1] = g.pv[k];
  return true;
}

extern "C" bool cl_release_device(void) {
  if (!g.up) return true;
  const int rc = zeta_shutdown();
  g.up = false;
  g.have_root = false;
  g.have_result = false;
  if (rc) return report("cl_release_device", rc);
  return true;
}
