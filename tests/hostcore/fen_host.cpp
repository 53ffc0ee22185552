// Host build of the CECP front end's board I/O (zeta_amd/host/zeta_engine.cpp)
// for CPU tests: FEN in (setboard), moves in coordinate notation (can2move,
// domove), FEN out (createfen).  The front end's main is renamed and never
// called; nothing here touches a device.
#define main zeta_engine_main
#include "../../zeta_amd/host/zeta_engine.cpp"
#undef main

#include <cstring>

// Sets up `fen`, plays the space-separated coordinate moves in `moves`, and
// writes the resulting FEN to out[cap].  Returns 0, or -1 for a FEN the
// parser rejects, -2 for a move that is not legal in its position, -3 for an
// output buffer that is too small.
extern "C" int fh_play(const char *fen, const char *moves, char *out, int cap) {
  uint64_t b[7];
  bool stm;
  int64_t gp = 0;
  if (!setboard(b, stm, fen, &gp)) return -1;
  int32_t ply = 0;
  const char *s = moves;
  while (s && *s) {
    while (*s == ' ') ++s;
    if (!*s) break;
    char mv[8] = {0};
    int k = 0;
    while (s[k] && s[k] != ' ' && k < 7) { mv[k] = s[k]; ++k; }
    s += k;
    const uint32_t m = can2move(mv, b, stm);
    if (!m) return -2;
    domove(b, stm, m);
    stm = !stm;
    ++ply;
  }
  const std::string f = createfen(b, stm, gp, ply);
  if ((int)f.size() + 1 > cap) return -3;
  std::memcpy(out, f.c_str(), f.size() + 1);
  return 0;
}

// Prints the board of `fen` as the front end's board print does (stdout).
// Returns 0, or -1 for a FEN the parser rejects.
extern "C" int fh_print(const char *fen) {
  uint64_t b[7];
  bool stm;
  int64_t gp = 0;
  if (!setboard(b, stm, fen, &gp)) return -1;
  printboard(b, stm, gp, 0);
  return 0;
}

// Position i of the front end's SMP benchmark sets (suite 0: benchkaufmann,
// 1: benchhyatt24), written back as a FEN; the number of legal moves in
// *nmoves.  Returns 0, -1 for a FEN the parser rejects, -2 for an index out of
// range, -3 for an output buffer that is too small.
extern "C" int fh_suite(int suite, int i, char *out, int cap, int *nmoves) {
  const char *const *set = suite ? HYATT24 : KAUFMANN;
  const int n = suite ? 24 : 25;
  if (i < 0 || i >= n) return -2;
  uint64_t b[7];
  bool stm;
  int64_t gp = 0;
  if (!setboard(b, stm, set[i], &gp)) return -1;
  *nmoves = (int)legal_moves(b, stm).size();
  const std::string f = createfen(b, stm, gp, 0);
  if ((int)f.size() + 1 > cap) return -3;
  std::memcpy(out, f.c_str(), f.size() + 1);
  return 0;
}

// The front end's clock: at game ply `ply` and a measured speed of `nps`
// nodes/s, runs the GUI command `line` (when not empty), then, when
// elapsed_ms >= 0, the update after an engine move that took elapsed_ms.
// Writes TimeBase, TimeLeft, TimeInc, MaxTime (ms), MovesLeft, MaxMoves,
// timemode and MaxNodes to out[8].  Returns 0, or -1 for a line that is not
// a clock command.
extern "C" int fh_clock(const char *line, int ply, long long nps, double elapsed_ms, double *out) {
  PLY = ply;
  nodes_per_second = nps;
  if (line && *line) {
    char cmd[64] = {0};
    sscanf(line, "%63s", cmd);
    if (!time_command(cmd, line)) return -1;
  }
  if (elapsed_ms >= 0) time_control(elapsed_ms);
  const double v[8] = {TimeBase, TimeLeft, TimeInc, MaxTime, (double)MovesLeft, (double)MaxMoves, (double)timemode,
                       (double)MaxNodes};
  for (int i = 0; i < 8; i++) out[i] = v[i];
  return 0;
}
