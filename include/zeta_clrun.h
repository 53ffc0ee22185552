/* zeta_clrun.h -- drop-in replacement of the reference's device boundary
 * (src/clrun.h:26-31) over libzetahip.so.
 *
 * The six functions keep the reference's names, signatures and call protocol
 * (bool return, false = failed; the reason is printed to stderr, where the
 * reference logged it with print_debug): a host front end written against
 * src/clrun.h links this shim instead of src/clrun.c and runs unchanged.
 * As in the reference, data passes through process globals that the host
 * program defines (src/zeta.h:48-134, src/zeta.c:56-110):
 *
 *   in:  GLOBAL_BOARD[7]  quad-bitboard of the root (src/types.h:59-66)
 *        GLOBAL_HASHHISTORY[totalWorkUnits * MAXGAMEPLY], PLY
 *        threadsX, threadsY, totalWorkUnits   workers (ABDADA) per root
 *        tt1_memory (MB)                      transposition table size
 *        opencl_device_id                     HIP device ordinal here
 *   out: COUNTERS[totalWorkUnits * 64]  [w*64+1] nodes, [3] TT move hits,
 *                                       [4] TT score hits, [5] IID hits (0)
 *        PV[MAXPLY]                     PV[0] = root score (as Score), PV[1..] = best move and
 *                                       the principal variation read from the TT
 *
 * Differences, by design: the kernel name passed to cl_init_device selects
 * nothing (one library serves perft and search); cl_run_alphabeta(stm, d,
 * nodes) searches with iterative deepening 1..d inside one call (earlier
 * iterations of the host's own loop hit the table exactly, so the repeated
 * work is small); the node budget `nodes` applies per worker, as in the
 * reference (src/zeta.cl:1945); of the positions before the root in
 * GLOBAL_HASHHISTORY (worker 0's copy, entries 0..PLY-1) the newest
 * ZETA_HISTORY (128) are consulted for repetitions, together with the
 * searched path (src/zeta.cl:1560-1577). */
#ifndef ZETA_CLRUN_H
#define ZETA_CLRUN_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* process globals, defined by the host program (src/zeta.c:56-110) */
extern uint64_t threadsX;
extern uint64_t threadsY;
extern const uint64_t threadsZ; /* 64 */
extern uint64_t totalWorkUnits;
extern uint64_t tt1_memory;
extern uint64_t tt2_memory;
extern int32_t opencl_device_id;
extern int32_t opencl_platform_id;
extern int32_t PLY;
extern uint64_t *GLOBAL_BOARD;
extern uint64_t *GLOBAL_HASHHISTORY;
extern uint64_t *COUNTERS;
extern uint32_t *PV;

/* The shim's own launch knobs (no counterpart in the reference; defined by
 * the shim, 0 = library default, so a host written against src/clrun.h
 * needs nothing more): zeta_search_opts.split_depth / .waves_per_simd, which
 * the front end reads from config.txt keys of the same names. */
extern int32_t zeta_split_depth;
extern int32_t zeta_waves_per_simd;

/* src/clrun.h:26 -- binds the device (zeta_init) once; later calls are no-ops */
bool cl_init_device(char *kernelname);
/* src/clrun.h:27 -- snapshots GLOBAL_BOARD and the hash history for the next run */
bool cl_write_objects(void);
/* src/clrun.h:28 -- fixed-depth alpha-beta + quiescence of the root */
bool cl_run_alphabeta(bool stm, int32_t depth, uint64_t nodes);
/* src/clrun.h:29 -- perft(depth) of the root */
bool cl_run_perft(bool stm, int32_t depth);
/* src/clrun.h:30 -- writes the last run's results to COUNTERS and PV */
bool cl_read_memory(void);
/* src/clrun.h:31 -- releases the device (zeta_shutdown) */
bool cl_release_device(void);

#ifdef __cplusplus
}
#endif
#endif
