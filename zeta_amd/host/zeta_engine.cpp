// zeta_engine.cpp -- CECP (xboard) front end of the MI355X build.
//
// The command set, the output strings and the config.txt keys follow the
// reference's front end (src/zeta.c:525-709, src/xboard.c:44-1499,
// src/io.c:335-582, src/search.c:35-271, src/test.c:43-140); the device work
// goes through the reference's six-function boundary (include/zeta_clrun.h),
// here backed by libzetahip.so.  Position handling (FEN, move text, make /
// unmake) uses the same chess core as the kernels (zeta_amd/csrc/chess.h,
// host build).
#include <ctype.h>
#include <getopt.h>
#include <inttypes.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/zeta_clrun.h"
#include "../../include/zeta_hip.h"
#include "../csrc/chess.h"

using zc::Legal;
using zc::Pos;

// ------------------------------------------------------- process globals
// (src/zeta.c:56-110), read and written by the clrun boundary
uint64_t threadsX = 1;
uint64_t threadsY = 1;
const uint64_t threadsZ = 64;
uint64_t totalWorkUnits = 1;
uint64_t tt1_memory = 0;
uint64_t tt2_memory = 0;
int32_t opencl_device_id = 0;
int32_t opencl_platform_id = 0;
int32_t PLY = 0;
int64_t GAMEPLY = 0;  // src/zeta.c:89: the game ply at the last setboard (2 * fullmove + stm)
uint64_t *GLOBAL_BOARD = nullptr;
uint64_t *GLOBAL_HASHHISTORY = nullptr;
uint64_t *COUNTERS = nullptr;
uint32_t *PV = nullptr;

namespace {

constexpr const char *VERSION = "099m-mi355x";
constexpr int MAXPLY = 64;         // src/types.h:76
constexpr int MAXGAMEPLY = 1024;   // src/types.h:77
constexpr int ESTEBF = 3;          // src/types.h:81
constexpr int32_t INF = 32000;     // src/types.h:87
constexpr int32_t MATESCORE = 30000;
constexpr const char *STARTFEN = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1";
constexpr const char *LOGFILE = "zeta.log";

// engine state (src/zeta.c:67-110)
uint64_t BOARD[7];
bool STM = false;
int32_t SD = MAXPLY;
uint64_t MaxNodes = 1;
int64_t nodes_per_second = 0;
// time control in milliseconds, 5 s per move until the GUI sets one
// (src/zeta.c:78-85); timemode 0 = fixed time per move, 1 = conventional
// clock (MaxMoves per session), 2 = base plus increment
double MaxTime = 5 * 1000, TimeBase = 5 * 1000, TimeInc = 0, TimeLeft = 5 * 1000;
int32_t MovesLeft = 1, MaxMoves = 1, timemode = 0;
uint64_t ABNODECOUNT = 0, TTHITS = 0, TTSCOREHITS = 0, IIDHITS = 0;
bool xboard_mode = false, xboard_force = false, xboard_post = false, xboard_debug = false;
int xboard_protover = 0;
FILE *LogFile = nullptr;
const char *configfile = "config.txt";

// game history: the position before each ply, for undo / remove
struct Undo {
  uint64_t b[7];
  bool stm;
};
std::vector<Undo> history;
std::vector<uint64_t> HashHistory(MAXGAMEPLY + 1, 0);

double now_ms() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

void logf_(const char *fmt, ...) {
  if (!LogFile) return;
  time_t t = time(nullptr);
  char d[64];
  strftime(d, sizeof d, "[%Y-%m-%d %H:%M:%S] ", localtime(&t));
  fputs(d, LogFile);
  va_list ap;
  va_start(ap, fmt);
  vfprintf(LogFile, fmt, ap);
  va_end(ap);
  fflush(LogFile);
}

// print to stdout and to the log
void out(const char *fmt, ...) {
  char buf[4096];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  fputs(buf, stdout);
  fflush(stdout);
  logf_("%s", buf);
}

[[noreturn]] void quitengine(int code) {
  (void)cl_release_device();
  free(GLOBAL_BOARD);
  free(GLOBAL_HASHHISTORY);
  free(COUNTERS);
  free(PV);
  if (LogFile) fclose(LogFile);
  exit(code);
}

// ------------------------------------------------------------ positions
Pos to_pos(const uint64_t *b, bool stm) {
  Pos p;
  p.b0 = b[0];
  p.b1 = b[1];
  p.b2 = b[2];
  p.b3 = b[3];
  p.cr = b[4];
  p.stm = stm ? 1u : 0u;
  return p;
}

std::vector<uint32_t> legal_moves(const uint64_t *b, bool stm) {
  const Pos p = to_pos(b, stm);
  Legal L;
  zc::analyse(p, L);
  std::vector<uint32_t> mv;
  uint64_t pcs = L.own;
  while (pcs) {
    const int f = zc::lsb(pcs);
    pcs &= pcs - 1;
    uint64_t t = zc::piece_targets(p, L, f, false);
    while (t) {
      const int to = zc::lsb(t);
      t &= t - 1;
      mv.push_back(zc::make_code(p, f, to));
    }
  }
  return mv;
}

bool in_check(const uint64_t *b, bool stm) {
  zc::Node0 B;
  zc::node_basics(to_pos(b, stm), B);
  return B.checkers != 0;
}

// host domove (src/bitboard.c:583-693): planes, moved flags with the castle
// rook, e.p. flag, half-move clock, hash from scratch
void domove(uint64_t *b, bool stm, uint32_t m) {
  if (!(m & 0x3FFFFFFF)) return;
  const Pos c = zc::apply(to_pos(b, stm), m, true);
  b[0] = c.b0;
  b[1] = c.b1;
  b[2] = c.b2;
  b[3] = c.b3;
  b[4] = c.cr;
  b[6] = zc::next_hmc((uint32_t)b[6], m);
  b[5] = zc::hash_full(c);
}

// FEN -> quad-bitboard with the reference's rules (src/io.c:335-532): moved
// flags start all set, castle letters clear rook and king squares, the FEN
// e.p. square is cleared on rank 3/6 where the move generator never looks.
// gameply (optional): the FEN's move number as the reference counts it,
// 2 * fullmove + stm (src/io.c:500)
bool setboard(uint64_t *b, bool &stm, const char *fen, int64_t *gameply = nullptr) {
  char position[1024] = {0}, cstm[16] = {0}, castle[16] = {0}, cep[16] = {0};
  unsigned long long hmc = 0, fullmove = 1;
  if (sscanf(fen, "%1023s %15s %15s %15s %llu %llu", position, cstm, castle, cep, &hmc, &fullmove) < 1)
    return false;
  uint64_t nb[7] = {0, 0, 0, 0, ~0ULL, 0, 0};
  int rank = 7, file = 0;
  for (const char *q = position; *q; q++) {
    const char ch = *q;
    if (ch == '/') {
      rank--;
      file = 0;
      if (rank < 0) return false;
      continue;
    }
    if (ch >= '1' && ch <= '8') {
      file += ch - '0';
      continue;
    }
    const char *types = "pnkbrq";  // piece type 1..6 (src/types.h:94-100)
    const char lower = (char)(ch | 0x20);
    const char *tp = strchr(types, lower);
    if (!tp || !*tp || file > 7) return false;
    const uint32_t type = (uint32_t)(tp - types) + 1;
    const uint32_t color = (ch == lower) ? 1u : 0u;
    const int sq = rank * 8 + file;
    const uint64_t m = 1ULL << sq;
    if (color) nb[0] |= m;
    if (type & 1) nb[1] |= m;
    if (type & 2) nb[2] |= m;
    if (type & 4) nb[3] |= m;
    file++;
  }
  if (rank != 0) return false;
  uint64_t cr = 0;
  if (castle[0] != '-')
    for (const char *q = castle; *q; q++) {
      if (*q == 'Q') cr |= zc::SMCRWHITEQ;
      if (*q == 'K') cr |= zc::SMCRWHITEK;
      if (*q == 'q') cr |= zc::SMCRBLACKQ;
      if (*q == 'k') cr |= zc::SMCRBLACKK;
    }
  nb[4] ^= cr;
  if (cep[0] >= 'a' && cep[0] <= 'h' && cep[1] >= '1' && cep[1] <= '8') {
    const int sq = (cep[1] - '1') * 8 + (cep[0] - 'a');
    if (sq > 0) nb[4] &= ~(1ULL << sq);
  }
  nb[6] = hmc;
  const bool nstm = (cstm[0] == 'b' || cstm[0] == 'B');
  // one king per side (src/bitboard.c:448-459)
  const uint64_t kings = nb[1] & nb[2] & ~nb[3];
  const uint64_t occ = nb[1] | nb[2] | nb[3];
  if (__builtin_popcountll(nb[0] & kings) != 1 || __builtin_popcountll((nb[0] ^ occ) & kings) != 1) return false;
  nb[5] = zc::hash_full(to_pos(nb, nstm));
  memcpy(b, nb, sizeof nb);
  stm = nstm;
  if (gameply) *gameply = (int64_t)fullmove * 2 + (nstm ? 1 : 0);
  return true;
}

std::string move2can(uint32_t m) {
  if (!(m & 0x3FFFFFFF)) return "0000";
  const int f = (int)zc::mv_from(m), t = (int)zc::mv_to(m);
  std::string s;
  s += (char)('a' + (f & 7));
  s += (char)('1' + (f >> 3));
  s += (char)('a' + (t & 7));
  s += (char)('1' + (t >> 3));
  const uint32_t pf = (zc::mv_pfrom(m) >> 1) & 7, pt = (zc::mv_pto(m) >> 1) & 7;
  if (pf == 1 && pt != 1) s += "?pnkbrq"[pt];
  return s;
}

// text -> legal move; 0 if illegal.  The engine's own promotions are to a
// queen (Q1), but a move the GUI sends may promote to a knight, bishop or
// rook, and the board must follow it (src/io.c:104-113); no letter or 'q'
// is a queen
uint32_t can2move(const char *s, const uint64_t *b, bool stm) {
  if (strlen(s) < 4) return 0;
  const int f = (s[0] - 'a') + 8 * (s[1] - '1'), t = (s[2] - 'a') + 8 * (s[3] - '1');
  for (uint32_t m : legal_moves(b, stm)) {
    if ((int)zc::mv_from(m) != f || (int)zc::mv_to(m) != t) continue;
    const bool promo = ((zc::mv_pfrom(m) >> 1) & 7) == zc::PAWN && ((zc::mv_pto(m) >> 1) & 7) != zc::PAWN;
    const char c = (char)tolower((unsigned char)s[4]);
    const int type = c == 'n' ? zc::KNIGHT : c == 'b' ? zc::BISHOP : c == 'r' ? zc::ROOK : 0;
    if (promo && type) m = (m & ~(0xFu << 22)) | ((uint32_t)((type << 1) | (stm ? 1 : 0)) << 22);
    return m;
  }
  return 0;
}

// the board as a FEN string (src/io.c:227-330): castle rights from the moved
// flags ("-" once every flagged square has moved), the e.p. target from the
// cleared flag on ranks 4/5, the move number as (gameply + ply) / 2
std::string createfen(const uint64_t *b, bool stm, int64_t gameply, int32_t ply) {
  static const char *wp = " PNKBRQ", *bp = " pnkbrq";
  std::string s;
  for (int r = 7; r >= 0; r--) {
    int spaces = 0;
    for (int f = 0; f < 8; f++) {
      const uint32_t pc = zc::piece_at(to_pos(b, false), r * 8 + f);
      const uint32_t type = (pc >> 1) & 7;
      if (type && spaces) {
        s += std::to_string(spaces);
        spaces = 0;
      }
      if (type) s += (pc & 1) ? bp[type] : wp[type];
      else spaces++;
    }
    if (spaces) s += std::to_string(spaces);
    if (r > 0) s += "/";
  }
  s += stm ? " b " : " w ";
  const uint64_t ncr = ~b[4];
  if ((ncr & zc::SMCRALL) == 0) {
    s += "-";
  } else {
    if ((ncr & zc::SMCRWHITEK) == zc::SMCRWHITEK) s += "K";
    if ((ncr & zc::SMCRWHITEQ) == zc::SMCRWHITEQ) s += "Q";
    if ((ncr & zc::SMCRBLACKK) == zc::SMCRBLACKK) s += "k";
    if ((ncr & zc::SMCRBLACKQ) == zc::SMCRBLACKQ) s += "q";
  }
  s += " ";
  const uint64_t ep = ncr & 0x000000FFFF000000ULL;
  int sq = ep ? __builtin_ctzll(ep) : 0;
  if (sq) {
    sq += stm ? -8 : 8;
    s += (char)('a' + (sq & 7));
    s += (char)('1' + (sq >> 3));
  } else {
    s += "-";
  }
  s += " " + std::to_string((unsigned long long)b[6]) + " " + std::to_string((long long)((gameply + ply) / 2));
  return s;
}

// src/io.c:164-226: the board, "-" for an empty square, then its FEN
void printboard(const uint64_t *b, bool stm, int64_t gameply, int32_t ply) {
  static const char *wp = "-PNKBRQ", *bp = "-pnkbrq";
  std::string s = "###ABCDEFGH###\n";
  for (int r = 7; r >= 0; r--) {
    s += "#" + std::to_string(r + 1) + " ";
    for (int f = 0; f < 8; f++) {
      const uint32_t pc = zc::piece_at(to_pos(b, false), r * 8 + f);
      const uint32_t type = (pc >> 1) & 7;
      s += (pc & 1) ? bp[type] : wp[type];
    }
    s += "\n";
  }
  s += "###ABCDEFGH###\n";
  s += "#fen: " + createfen(b, stm, gameply, ply) + "\n";
  out("%s", s.c_str());
}

// --------------------------------------------------------- host drivers
void fill_history() {
  for (uint64_t w = 0; w < totalWorkUnits; w++)
    for (int i = 0; i <= PLY && i < MAXGAMEPLY; i++) GLOBAL_HASHHISTORY[w * MAXGAMEPLY + i] = HashHistory[i];
}

// src/search.c:225-271
void perft(const uint64_t *board, bool stm, int32_t depth) {
  ABNODECOUNT = 0;
  memcpy(GLOBAL_BOARD, board, 7 * sizeof(uint64_t));
  fill_history();
  if (!cl_write_objects() || !cl_run_perft(stm, depth) || !cl_read_memory()) quitengine(EXIT_FAILURE);
  for (uint64_t i = 0; i < totalWorkUnits; i++) ABNODECOUNT += COUNTERS[i * 64 + 1];
}

bool is_mate(int32_t v) { return (v > MATESCORE && v < INF) || (v < -MATESCORE && v > -INF); }

// src/search.c:35-223: iterative deepening over single-depth device calls
uint32_t rootsearch(const uint64_t *board, bool stm, int32_t depth) {
  uint32_t bestmove = 0;
  int32_t bestscore = 0;
  ABNODECOUNT = TTHITS = TTSCOREHITS = IIDHITS = 0;
  const double start = now_ms();
  memcpy(GLOBAL_BOARD, board, 7 * sizeof(uint64_t));
  fill_history();
  if (!xboard_mode) out("depth score time nodes pv \n");
  int32_t idf = 1;
  double elapsed = 0;
  do {
    if (!cl_write_objects() || !cl_run_alphabeta(stm, idf, MaxNodes / totalWorkUnits) || !cl_read_memory())
      quitengine(EXIT_FAILURE);
    for (uint64_t i = 0; i < totalWorkUnits; i++) {
      ABNODECOUNT += COUNTERS[i * 64 + 1];
      TTHITS += COUNTERS[i * 64 + 3];
      TTSCOREHITS += COUNTERS[i * 64 + 4];
      IIDHITS += COUNTERS[i * 64 + 5];
    }
    elapsed = (now_ms() - start + 1) / 1000;
    const uint32_t pvmove = PV[1] & 0x3FFFFFFF;
    if (idf == 1 && pvmove) bestmove = PV[1];
    if (COUNTERS[1] < MaxNodes / totalWorkUnits) {
      if (pvmove) bestmove = PV[1];
      const int32_t s = (int32_t)PV[0];
      bestscore = (s == INF || s == -INF) ? 0 : s;
      int32_t xs = bestscore;
      if (bestscore <= -MATESCORE) xs = -100000 - (INF + bestscore);
      if (bestscore >= MATESCORE) xs = 100000 - (-INF + bestscore);
      if ((xboard_post || !xboard_mode) && pvmove) {
        // thinking line with the principal variation (src/search.c:150-168)
        std::string line;
        for (int i = 1; i < MAXPLY && i <= idf; i++) {
          if (!(PV[i] & 0x3FFFFFFF)) break;
          line += move2can(PV[i]) + " ";
        }
        out("%i %i %i %" PRIu64 " %s\n", idf, xs, (int32_t)(elapsed * 100), ABNODECOUNT, line.c_str());
      } else {
        // as the reference (src/search.c:170-171): no thinking line ends the
        // deepening, so xboard mode without post plays the first completed
        // iteration's move
        break;
      }
    }
    if (is_mate(bestscore) && bestscore >= MATESCORE && idf >= INF - bestscore) break;
    if (is_mate(bestscore) && bestscore <= -MATESCORE && idf >= INF + bestscore) break;
    idf++;
  } while (idf <= depth && elapsed * ESTEBF < MaxTime && ABNODECOUNT * ESTEBF <= MaxNodes && ABNODECOUNT > 1 &&
           idf < MAXPLY);
  if (!xboard_mode || xboard_debug)
    out("#%" PRIu64 " searched nodes in %lf seconds, with %" PRIu64 " ttmovehits, and %" PRIu64
        " ttscorehits, %" PRIu64 " iidhits, ebf: %lf, nps: %" PRIu64 " \n",
        ABNODECOUNT, elapsed, TTHITS, TTSCOREHITS, IIDHITS, pow((double)ABNODECOUNT, 1.0 / idf),
        (uint64_t)((double)ABNODECOUNT / elapsed));
  // nps estimate for time control (src/search.c:214-220)
  const int64_t nps = (int64_t)((double)ABNODECOUNT / elapsed);
  nodes_per_second += (nps > nodes_per_second) ? (int64_t)((nps - nodes_per_second) * 0.66)
                                               : (int64_t)((nps - nodes_per_second) * 0.33);
  return bestmove;
}

// ----------------------------------------------------------- game flow
void push_move(uint32_t m) {
  history.push_back({{BOARD[0], BOARD[1], BOARD[2], BOARD[3], BOARD[4], BOARD[5], BOARD[6]}, STM});
  domove(BOARD, STM, m);
  PLY++;
  STM = !STM;
  if (PLY <= MAXGAMEPLY) HashHistory[PLY] = BOARD[5];
}

void pop_move() {
  if (history.empty()) return;
  const Undo &u = history.back();
  memcpy(BOARD, u.b, sizeof BOARD);
  STM = u.stm;
  history.pop_back();
  PLY--;
}

void new_game(const char *fen) {
  bool stm = false;
  uint64_t b[7];
  int64_t gp = 0;
  if (!setboard(b, stm, fen, &gp)) {
    out("Error (in setting chess psotition via fen string): setboard\n");
    out("tellusererror (Error in setting chess psotition via fen string): setboard\n");
    return;
  }
  memcpy(BOARD, b, sizeof BOARD);
  STM = stm;
  GAMEPLY = gp;
  PLY = 0;
  history.clear();
  HashHistory.assign(MAXGAMEPLY + 1, 0);
  HashHistory[0] = BOARD[5];
}

// moves left in the current session of a conventional clock, one move of
// time kept spare (src/xboard.c:416-417)
int32_t session_moves_left() { return (MaxMoves - (((PLY + 1) / 2) % MaxMoves)) + 1; }

// the node budget of ms milliseconds at the measured speed (0 once the
// clock has run out)
uint64_t nodes_for(double ms) {
  const double n = ms / 1000 * (double)nodes_per_second;
  return n > 0 ? (n < 1.8e19 ? (uint64_t)n : UINT64_MAX) : 0;
}

// the budget of the next move and its node count
void set_move_budget() {
  MaxTime = TimeLeft / MovesLeft + TimeInc;
  MaxNodes = nodes_for(MaxTime);
}

// the GUI's clock commands (src/xboard.c:384-455):
//   level MPS BASE INC  BASE in minutes or minutes:seconds, INC in seconds;
//                       MPS 0 = base plus increment (40 moves assumed),
//                       else MPS moves per session
//   st SECONDS          a fixed time per move
//   time CENTISECONDS   the engine's clock
// Returns false for any other command.  A level line in neither form leaves
// the clock as it was (the reference has cleared it by then).
bool time_command(const char *cmd, const char *line) {
  if (!strcmp(cmd, "level")) {
    int mps = 0, mins = 0, secs = 0;
    double inc = 0;
    if (sscanf(line, "level %d %d %lf", &mps, &mins, &inc) != 3) {
      secs = 0;
      inc = 0;
      if (sscanf(line, "level %d %d:%d %lf", &mps, &mins, &secs, &inc) != 4) return true;
    }
    if (mps < 0) return true;
    MaxMoves = mps;
    TimeBase = (60.0 * mins + secs) * 1000;
    TimeInc = inc * 1000;
    TimeLeft = TimeBase;
    timemode = MaxMoves == 0 ? 2 : 1;
    MovesLeft = timemode == 2 ? 40 : session_moves_left();
    set_move_budget();
    return true;
  }
  if (!strcmp(cmd, "st")) {
    double st = 0;
    sscanf(line, "st %lf", &st);
    TimeBase = TimeLeft = st * 1000;
    TimeInc = 0;
    MovesLeft = MaxMoves = 1;
    timemode = 0;
    set_move_budget();
    return true;
  }
  if (!strcmp(cmd, "time")) {
    double cs = 0;
    sscanf(line, "time %lf", &cs);
    TimeLeft = cs * 10;
    if (timemode == 1) MovesLeft = session_moves_left();
    set_move_budget();
    return true;
  }
  return false;
}

// after the engine's move (src/xboard.c:355-371)
void time_control(double elapsed_ms) {
  TimeLeft -= elapsed_ms;
  if (timemode == 1) MovesLeft = session_moves_left();
  if (timemode == 0) TimeLeft = TimeBase;
  if (timemode == 2) TimeLeft += TimeInc;
  MaxTime = TimeLeft / MovesLeft + TimeInc;
  if (timemode == 1 && MovesLeft == 2) TimeLeft += TimeBase;
  MaxNodes = nodes_for(MaxTime);
}

// think and play a move (src/xboard.c:217-376)
void think() {
  const bool kic = in_check(BOARD, STM);
  const double start = now_ms();
  HashHistory[PLY] = BOARD[5];
  if (PLY >= MAXGAMEPLY) {
    out(STM ? "result 1-0 { resign - max game ply reached }\n" : "result 0-1 { resign - max game ply reached}\n");
    return;
  }
  uint32_t move = rootsearch(BOARD, STM, SD);
  const bool none = !(move & 0x3FFFFFFF);
  const std::vector<uint32_t> legal = legal_moves(BOARD, STM);
  if (ABNODECOUNT == 0 && !none) {
    out(STM ? "result 1-0 { resign - internal error }\n" : "result 0-1 { resign - internal error}\n");
  } else if (legal.empty() && kic) {
    out(STM ? "result 1-0 { checkmate }\n" : "result 0-1 { checkmate }\n");
  } else if (legal.empty()) {
    out("result 1/2-1/2 { stalemate }\n");
  } else {
    // a search that returned no move while moves exist (none does: the depth-1
    // iteration always completes) must not put a null move on the board
    if (none) move = legal[0];
    out("move %s\n", move2can(move).c_str());
    push_move(move);
    if (!xboard_mode || xboard_debug) printboard(BOARD, STM, GAMEPLY, PLY);
    time_control(now_ms() - start);
  }
}

// one fixed-depth search of a position with the configured workers, the
// reference's benchmark() (src/bench.c:36-112); returns the seconds taken
double benchmark(const uint64_t *board, bool stm, int32_t depth) {
  ABNODECOUNT = 0;
  TTHITS = 0;
  memcpy(GLOBAL_BOARD, board, 7 * sizeof(uint64_t));
  fill_history();
  const double start = now_ms();
  if (!cl_write_objects() || !cl_run_alphabeta(stm, depth, MaxNodes) || !cl_read_memory()) return -1;
  const double el = (now_ms() - start + 1) / 1000;
  for (uint64_t i = 0; i < totalWorkUnits; i++) {
    ABNODECOUNT += COUNTERS[i * 64 + 1];
    TTHITS += COUNTERS[i * 64 + 3];
  }
  const int32_t sc = (int32_t)PV[0];
  out("depth: %i, nodes %" PRIu64 ", nps: %i, time: %lf sec, score: %i  move %s\n", depth, ABNODECOUNT,
      (int)((double)ABNODECOUNT / el), el, (sc == INF || sc == -INF) ? 0 : sc, move2can(PV[1]).c_str());
  return el;
}

// ---------------------------------------------------------------- tests
// the reference's known-answer perft positions (src/test.c:43-73)
struct Kat {
  const char *fen;
  int depth;
  uint64_t nodes;
};
const Kat KATS[] = {
    {"rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq -", 1, 20},
    {"rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq -", 2, 400},
    {"rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq -", 3, 8902},
    {"rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq -", 4, 197281},
    {"8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - ", 1, 14},
    {"8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - ", 2, 191},
    {"8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - ", 3, 2812},
    {"8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - ", 4, 43238},
    {"r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - ", 1, 48},
    {"r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - ", 2, 2039},
    {"r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - ", 3, 97862},
    {"r4rk1/1pp1qppp/p1np1n2/2b1p1B1/2B1P1b1/P1NP1N2/1PP1QPPP/R4RK1 w - -", 1, 46},
    {"r4rk1/1pp1qppp/p1np1n2/2b1p1B1/2B1P1b1/P1NP1N2/1PP1QPPP/R4RK1 w - -", 2, 2079},
    {"r4rk1/1pp1qppp/p1np1n2/2b1p1B1/2B1P1b1/P1NP1N2/1PP1QPPP/R4RK1 w - -", 3, 89890},
};

// src/test.c:43-140
bool selftest() {
  const uint64_t todo = sizeof KATS / sizeof KATS[0];
  uint64_t passed = 0;
  uint64_t b[7];
  bool stm = false;
  int64_t gp = 0;
  for (uint64_t done = 0; done < todo; done++) {
    out("#> doing perft depth: %d for position %" PRIu64 " of %" PRIu64 "\n", KATS[done].depth, done + 1, todo);
    if (!setboard(b, stm, KATS[done].fen, &gp)) {
      out("# Error (in setting fen position): setboard\n");
      continue;
    }
    printboard(b, stm, gp, 0);
    const double start = now_ms();
    perft(b, stm, KATS[done].depth);
    const double el = (now_ms() - start) / 1000;
    if (ABNODECOUNT == KATS[done].nodes) {
      passed++;
      out("#> OK, Nodecount Correct, %" PRIu64 " nodes in %lf seconds with %" PRIu64 " nps.\n", ABNODECOUNT, el,
          (uint64_t)(ABNODECOUNT / (el > 0 ? el : 1e-9)));
    } else {
      out("#> Error, Nodecount NOT Correct, %" PRIu64 " computed nodes != %" PRIu64
          " nodes for depth %d. in %lf seconds with %" PRIu64 " nps.\n",
          ABNODECOUNT, KATS[done].nodes, KATS[done].depth, el, (uint64_t)(ABNODECOUNT / (el > 0 ? el : 1e-9)));
    }
  }
  // the reference's summary banner (src/test.c:144-146)
  out("#\n###############################\n");
  out("### passed %" PRIu64 " from %" PRIu64 " tests ###\n", passed, todo);
  out("###############################\n");
  return passed == todo;
}

// ---------------------------------------------------------------- config
// src/io.c:533-582
bool read_and_init_config(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) {
    out("Error (%s file missing): try --guessconfigx option to create a config.txt file or --help option for "
        "further options\n", path);
    return false;
  }
  char line[1024];
  while (fgets(line, sizeof line, f)) {
    sscanf(line, "threadsX: %" SCNu64 ";", &threadsX);
    sscanf(line, "threadsY: %" SCNu64 ";", &threadsY);
    sscanf(line, "nodes_per_second: %" SCNi64 ";", &nodes_per_second);
    sscanf(line, "tt1_memory: %" SCNu64 ";", &tt1_memory);
    sscanf(line, "tt2_memory: %" SCNu64 ";", &tt2_memory);
    sscanf(line, "opencl_platform_id: %d;", &opencl_platform_id);
    sscanf(line, "opencl_device_id: %d;", &opencl_device_id);
    // this build's launch knobs (written by --guessconfig; absent = defaults)
    sscanf(line, "split_depth: %d;", &zeta_split_depth);
    sscanf(line, "waves_per_simd: %d;", &zeta_waves_per_simd);
  }
  fclose(f);
  MaxNodes = (uint64_t)nodes_per_second;
  totalWorkUnits = threadsX * threadsY;
  if (totalWorkUnits < 1) totalWorkUnits = 1;
  return true;
}

bool engineinits() {
  GLOBAL_BOARD = (uint64_t *)calloc(7, sizeof(uint64_t));
  GLOBAL_HASHHISTORY = (uint64_t *)calloc(totalWorkUnits * MAXGAMEPLY, sizeof(uint64_t));
  COUNTERS = (uint64_t *)calloc(totalWorkUnits * threadsZ, sizeof(uint64_t));
  PV = (uint32_t *)calloc(MAXPLY, sizeof(uint32_t));
  return GLOBAL_BOARD && GLOBAL_HASHHISTORY && COUNTERS && PV;
}

// Autotuner (replaces --guessconfig[x], src/clconfig.c:878-960, which
// doubles threadsY while the measured nodes per second still rise by a
// margin).  Coordinate descent over the search's launch knobs on seeded
// playout positions, each knob keeping the value with the shortest total
// time to depth (more workers always raise nodes/s, not always the time to
// depth):
//   1. workers per root (threadsX): each position searched alone, as the
//      engine searches (ID 1..depth, TT cleared per root);
//   2. ABDADA split depth (split_depth) at the chosen worker count;
//   3. persistent wavefronts per SIMD (waves_per_simd) on a batch of roots
//      large enough to fill the device (a single root never does).
// Writes config.txt with the winners, the TT size and the measured nodes per
// second, then reads it back.
bool guess_config(bool extended) {
  zeta_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.device = opencl_device_id;
  cfg.tt_log2 = 26;  // tt1_memory 1024 MB
  int rc = zeta_init(&cfg);
  auto fail = [&](const char *what) {
    char msg[512];
    zeta_last_error(msg, sizeof msg);
    out("Error (%s: %s): --guessconfig\n", what, msg);
    zeta_shutdown();
    return false;
  };
  if (rc) return fail("zeta_init");
  const int npos = extended ? 16 : 4, depth = extended ? 6 : 5;
  std::vector<zeta_pos> pos(npos);
  if ((rc = zeta_playouts(3, npos, 20, 60, pos.data()))) return fail("zeta_playouts");
  std::vector<zeta_result> res(npos);
  auto opts_of = [&](int w, int sd, int wps, int d) {
    zeta_search_opts o;
    memset(&o, 0, sizeof o);
    o.depth = d;
    o.workers = w;
    o.tt_log2 = 26;
    o.clear_tt = 1;
    o.split_depth = sd;
    o.waves_per_simd = wps;
    return o;
  };
  // one root at a time; total seconds and nodes/s
  auto single = [&](int w, int sd, double &secs, double &nps) {
    const zeta_search_opts o = opts_of(w, sd, 0, depth);
    secs = 0;
    uint64_t nodes = 0;
    for (int i = 0; i < npos; i++) {
      const double t0 = now_ms();
      if ((rc = zeta_search_batch(&pos[i], 1, &o, &res[i]))) return false;
      secs += (now_ms() - t0) / 1000;
      nodes += res[i].nodes;
    }
    nps = (double)nodes / (secs > 0 ? secs : 1e-9);
    return true;
  };
  int best_w = 1, best_sd = 2, best_wps = 0;
  double best_t = 0, best_nps = 0;
  const std::vector<int> wc =
      extended ? std::vector<int>{1, 2, 4, 8, 16, 32, 64, 128, 256} : std::vector<int>{1, 4, 16, 64, 128};
  for (int w : wc) {
    double t, nps;
    if (!single(w, best_sd, t, nps)) return fail("workers");
    out("#> workers per root %d: time to depth %d over %d positions %.3f s, %.0f nodes/s\n", w, depth, npos, t, nps);
    if (best_t == 0 || t < best_t) best_t = t, best_nps = nps, best_w = w;
  }
  if (best_w > 1) {
    const std::vector<int> sc = extended ? std::vector<int>{1, 2, 3, 4, 5} : std::vector<int>{1, 2, 3};
    for (int sd : sc) {
      if (sd == 2) continue;  // measured above
      double t, nps;
      if (!single(best_w, sd, t, nps)) return fail("split_depth");
      out("#> split depth %d (workers %d): time to depth %d over %d positions %.3f s, %.0f nodes/s\n", sd, best_w,
          depth, npos, t, nps);
      if (t < best_t) best_t = t, best_nps = nps, best_sd = sd;
    }
  }
  {
    // a batch of roots that asks for more wavefronts than the device holds
    const int nb = extended ? 1024 : 256;
    std::vector<zeta_pos> bp(nb);
    std::vector<zeta_result> br(nb);
    if ((rc = zeta_playouts(5, nb, 20, 60, bp.data()))) return fail("zeta_playouts");
    double bt = 0;
    for (int wps : {2, 1}) {
      const zeta_search_opts o = opts_of(best_w < 8 ? 8 : best_w, best_sd, wps, depth - 1);
      const double t0 = now_ms();
      if ((rc = zeta_search_batch(bp.data(), nb, &o, br.data()))) return fail("waves_per_simd");
      const double t = (now_ms() - t0) / 1000;
      out("#> waves per SIMD %d: batch of %d roots to depth %d in %.3f s\n", wps, nb, depth - 1, t);
      if (bt == 0 || t < bt) bt = t, best_wps = wps;
    }
  }
  zeta_shutdown();
  FILE *f = fopen(configfile, "w");
  if (!f) {
    out("Error (cannot write %s): --guessconfig\n", configfile);
    return false;
  }
  fprintf(f, "threadsX: %d;\nthreadsY: 1;\nnodes_per_second: %" PRIu64 ";\ntt1_memory: 1024;\ntt2_memory: 0;\n"
             "opencl_platform_id: 0;\nopencl_device_id: %d;\nsplit_depth: %d;\nwaves_per_simd: %d;\n",
          best_w, (uint64_t)best_nps, opencl_device_id, best_sd, best_wps);
  fprintf(f, "# threadsX          workers (ABDADA wavefronts) per root\n"
             "# threadsY          multiplier of threadsX (1)\n"
             "# nodes_per_second  measured search speed, seeds the time control\n"
             "# tt1_memory        transposition table in MB (2^n x 16 B entries)\n"
             "# tt2_memory        unused (ABDADA keeps its busy table beside the TT)\n"
             "# opencl_device_id  HIP device ordinal\n"
             "# split_depth       ABDADA defers busy siblings at nodes with this much depth left\n"
             "# waves_per_simd    persistent search wavefronts per SIMD\n");
  fclose(f);
  out("#> config written to %s (threadsX %d, split_depth %d, waves_per_simd %d, nodes_per_second %" PRIu64 ")\n",
      configfile, best_w, best_sd, best_wps, (uint64_t)best_nps);
  // read back what a later start will read
  threadsX = 0;
  zeta_split_depth = zeta_waves_per_simd = -1;
  if (!read_and_init_config(configfile)) return false;
  out("#> config read back: threadsX %" PRIu64 ", split_depth %d, waves_per_simd %d, nodes_per_second %" PRIi64
      "\n", threadsX, zeta_split_depth, zeta_waves_per_simd, nodes_per_second);
  return threadsX == (uint64_t)best_w && zeta_split_depth == best_sd && zeta_waves_per_simd == best_wps;
}

// The two position sets of the reference's SMP benchmarks (the same
// positions, so that the commands measure the same thing):
// benchkaufmann, src/xboard.c:986-1013
const char *const KAUFMANN[25] = {
    "1rbq1rk1/p1b1nppp/1p2p3/8/1B1pN3/P2B4/1P3PPP/2RQ1R1K w - -",
    "3r2k1/p2r1p1p/1p2p1p1/q4n2/3P4/PQ5P/1P1RNPP1/3R2K1 b - -",
    "3r2k1/1p3ppp/2pq4/p1n5/P6P/1P6/1PB2QP1/1K2R3 w - -",
    "r1b1r1k1/1ppn1p1p/3pnqp1/8/p1P1P3/5P2/PbNQNBPP/1R2RB1K w - -",
    "2r4k/pB4bp/1p4p1/6q1/1P1n4/2N5/P4PPP/2R1Q1K1 b - -",
    "r5k1/3n1ppp/1p6/3p1p2/3P1B2/r3P2P/PR3PP1/2R3K1 b - -",
    "2r2rk1/1bqnbpp1/1p1ppn1p/pP6/N1P1P3/P2B1N1P/1B2QPP1/R2R2K1 b - -",
    "5r1k/6pp/1n2Q3/4p3/8/7P/PP4PK/R1B1q3 b - -",
    "r3k2r/pbn2ppp/8/1P1pP3/P1qP4/5B2/3Q1PPP/R3K2R w KQkq -",
    "3r2k1/ppq2pp1/4p2p/3n3P/3N2P1/2P5/PP2QP2/K2R4 b - -",
    "q3rn1k/2QR4/pp2pp2/8/P1P5/1P4N1/6n1/6K1 w - -",
    "6k1/p3q2p/1nr3pB/8/3Q1P2/6P1/PP5P/3R2K1 b - -",
    "1r4k1/7p/5np1/3p3n/8/2NB4/7P/3N1RK1 w - -",
    "1r2r1k1/p4p1p/6pB/q7/8/3Q2P1/PbP2PKP/1R3R2 w - -",
    "r2q1r1k/pb3p1p/2n1p2Q/5p2/8/3B2N1/PP3PPP/R3R1K1 w - -",
    "8/4p3/p2p4/2pP4/2P1P3/1P4k1/1P1K4/8 w - -",
    "1r1q1rk1/p1p2pbp/2pp1np1/6B1/4P3/2NQ4/PPP2PPP/3R1RK1 w - -",
    "q4rk1/1n1Qbppp/2p5/1p2p3/1P2P3/2P4P/6P1/2B1NRK1 b - -",
    "r2q1r1k/1b1nN2p/pp3pp1/8/Q7/PP5P/1BP2RPN/7K w - -",
    "8/5p2/pk2p3/4P2p/2b1pP1P/P3P2B/8/7K w - -",
    "8/2k5/4p3/1nb2p2/2K5/8/6B1/8 w - -",
    "1B1b4/7K/1p6/1k6/8/8/8/8 w - -",
    "rn1q1rk1/1b2bppp/1pn1p3/p2pP3/3P4/P2BBN1P/1P1N1PP1/R2Q1RK1 b - -",
    "8/p1ppk1p1/2n2p2/8/4B3/2P1KPP1/1P5P/8 w - -",
    "8/3nk3/3pp3/1B6/8/3PPP2/4K3/8 w - -"};
// benchhyatt24, src/xboard.c:1214-1240
const char *const HYATT24[24] = {
    "r2qkbnr/ppp2p1p/2n5/3P4/2BP1pb1/2N2p2/PPPQ2PP/R1B2RK1 b kq -",
    "r2qkbnr/ppp2p1p/8/nB1P4/3P1pb1/2N2p2/PPPQ2PP/R1B2RK1 b kq -",
    "r2qkbnr/pp3p1p/2p5/nB1P4/3P1Qb1/2N2p2/PPP3PP/R1B2RK1 b kq -",
    "r2qkb1r/pp3p1p/2p2n2/nB1P4/3P1Qb1/2N2p2/PPP3PP/R1B1R1K1 b kq -",
    "r2q1b1r/pp1k1p1p/2P2n2/nB6/3P1Qb1/2N2p2/PPP3PP/R1B1R1K1 b - -",
    "r2q1b1r/p2k1p1p/2p2n2/nB6/3PNQb1/5p2/PPP3PP/R1B1R1K1 b - -",
    "r2q1b1r/p2k1p1p/2p5/nB6/3Pn1Q1/5p2/PPP3PP/R1B1R1K1 b - -",
    "r2q1b1r/p1k2p1p/2p5/nB6/3PR1Q1/5p2/PPP3PP/R1B3K1 b - -",
    "r2q1b1r/p1k2p1p/8/np6/3PR3/5Q2/PPP3PP/R1B3K1 b - -",
    "r4b1r/p1kq1p1p/8/np6/3P1R2/5Q2/PPP3PP/R1B3K1 b - -",
    "r6r/p1kqbR1p/8/np6/3P4/5Q2/PPP3PP/R1B3K1 b - -",
    "5r1r/p1kqbR1p/8/np6/3P1B2/5Q2/PPP3PP/R5K1 b - -",
    "5r1r/p2qbR1p/1k6/np2B3/3P4/5Q2/PPP3PP/R5K1 b - -",
    "5rr1/p2qbR1p/1k6/np2B3/3P4/2P2Q2/PP4PP/R5K1 b - -",
    "5rr1/p2qbR1p/1kn5/1p2B3/3P4/2P2Q2/PP4PP/4R1K1 b - -",
    "4qRr1/p3b2p/1kn5/1p2B3/3P4/2P2Q2/PP4PP/4R1K1 b - -",
    "5qr1/p3b2p/1kn5/1p1QB3/3P4/2P5/PP4PP/4R1K1 b - -",
    "5q2/p3b2p/1kn5/1p1QB1r1/P2P4/2P5/1P4PP/4R1K1 b - -",
    "5q2/p3b2p/1kn5/3QB1r1/p1PP4/8/1P4PP/4R1K1 b - -",
    "5q2/p3b2p/1k6/3QR1r1/p1PP4/8/1P4PP/6K1 b - -",
    "5q2/p3b2p/1k6/4Q3/p1PP4/8/1P4PP/6K1 b - -",
    "3q4/p3b2p/1k6/2P1Q3/p2P4/8/1P4PP/6K1 b - -",
    "3q4/p3b2p/8/1kP5/p2P4/8/1P2Q1PP/6K1 b - -",
    "3q4/p3b2p/8/2P5/pk1P4/3Q4/1P4PP/6K1 b - -"};

// The current position searched by rootsearch to depth SD with 1, 2, 4, ...
// workers up to the configured count, each count announced by the progress
// lines of src/xboard.c:846-876, then the table of nps, time to depth and
// their speedups (src/xboard.c:924-935).  Returns the times; the worker
// counts in `workers`.
std::vector<double> smp_runs(const char *name, std::vector<uint64_t> &workers) {
  const uint64_t wmax = totalWorkUnits;
  std::vector<double> t;
  std::vector<uint64_t> nps;
  workers.clear();
  for (uint64_t w = 1;; w *= 2) {
    if (w > wmax) w = wmax;
    out("### doing inits for %s depth %d: ###\n", name, SD);
    totalWorkUnits = w;
    if (!cl_release_device() || !cl_init_device((char *)"alphabeta_gpu")) quitengine(EXIT_FAILURE);
    out("### computing %s depth %d: ###\n", name, SD);
    out("### work-groups: %" PRIu64 " ###\n", w);
    const double start = now_ms();
    rootsearch(BOARD, STM, SD);
    const double el = (now_ms() - start + 1) / 1000;
    t.push_back(el);
    nps.push_back((uint64_t)((double)ABNODECOUNT / el));
    workers.push_back(w);
    if (w == wmax) break;
  }
  totalWorkUnits = wmax;
  out("### workers\t#nps\t\t#nps speedup\t#time in s\t#ttd speedup \t#relative ttd speedup ###\n");
  for (size_t i = 0; i < t.size(); i++) {
    const double npsx = i ? (double)nps[i] / (double)(nps[0] ? nps[0] : 1) : 1.0;
    out("### %" PRIu64 "\t\t%" PRIu64 "\t\t%lf\t%lf\t%lf\t%lf \n", workers[i], nps[i], npsx, t[i], t[0] / t[i],
        i ? t[i - 1] / t[i] : 1.0);
  }
  return t;
}

// SMP benchmark over a position set, benchkaufmann / benchhyatt24
// (src/xboard.c:975-1201 and 1203-1427): smp_runs on every position of the
// set, then the time-to-depth speedups over the summed times of all positions
void bench_suite(const char *name, const char *const *fens, int npos) {
  std::vector<double> tall;
  std::vector<uint64_t> wall;
  for (int posi = 0; posi < npos; posi++) {
    out("#\n#\n### setting up board %d: ###\n", posi + 1);
    new_game(fens[posi]);
    printboard(BOARD, STM, GAMEPLY, PLY);
    const std::vector<double> t = smp_runs(name, wall);
    if (tall.size() < t.size()) tall.resize(t.size(), 0.0);
    for (size_t i = 0; i < t.size(); i++) tall[i] += t[i];
  }
  out("#\n# overall results\n### workers\t#ttd speedup\t#rel ttd speedup ###\n");
  for (size_t i = 0; i < tall.size(); i++)
    out("### %" PRIu64 "\t\t%lf\t%lf \n", wall[i], tall[0] / tall[i], i ? tall[i - 1] / tall[i] : 1.0);
  if (!cl_release_device() || !cl_init_device((char *)"alphabeta_gpu")) quitengine(EXIT_FAILURE);
}

void print_help() {
  out("Zeta %s, MI355X build. Options:\n", VERSION);
  out(" -h, --help         Print help\n -v, --version      Print version\n");
  out(" -l, --log          Write output/debug to file zeta.log\n -s, --selftest     Run an internal test\n");
  out(" --guessconfig      Measure the device and write config.txt (quick)\n");
  out(" --guessconfigx     Measure the device and write config.txt (longer)\n");
  out(" -d 0               Set device ID to 0 for guessconfig (the HIP device; -p is read, unused)\n");
  out(" -pl, -dl           Say that there are no OpenCL platforms or devices to list\n");
  out(" --config FILE      Read FILE instead of ./config.txt\n");
  out("Commands outside xboard mode: sd N, setboard FEN, go, perft, selftest, bench, benchsmp,\n"
      " benchkaufmann, benchhyatt24 (search depth N with 1, 2, 4, ... workers), help, quit\n");
}

// --------------------------------------------------------------- xboard
void xboard() {
  char line[1024], cmd[1024];
  while (fgets(line, sizeof line, stdin)) {
    if (line[0] == '\n') continue;
    logf_(">> %s", line);
    cmd[0] = 0;
    sscanf(line, "%1023s", cmd);
    if (!strcmp(cmd, "xboard")) {
      out("feature done=0\n");
      xboard_mode = true;
    } else if (!strcmp(cmd, "protover")) {
      sscanf(line, "protover %d", &xboard_protover);
      if (xboard_mode && xboard_protover < 2) {
        out("Error (unsupported protocoll version,  < v2): protover\n");
        out("tellusererror (unsupported protocoll version, < v2): protover\n");
      } else {
        out("feature myname=\"Zeta %s\"\n", VERSION);
        out("feature ping=1\nfeature setboard=1\nfeature playother=0\nfeature san=0\nfeature usermove=1\n");
        out("feature time=1\nfeature draw=0\nfeature sigint=0\nfeature reuse=1\nfeature analyze=0\n");
        out("feature variants=\"normal\"\nfeature colors=0\nfeature ics=0\nfeature name=0\nfeature pause=0\n");
        out("feature nps=0\nfeature debug=1\nfeature memory=1\nfeature smp=0\nfeature exclude=0\nfeature done=1\n");
      }
    } else if (!strcmp(cmd, "accepted") || !strcmp(cmd, "rejected")) {
      if (strstr(line, "debug") && !strcmp(cmd, "accepted")) xboard_debug = true;
      // moves arrive only as usermove: a GUI that refuses it cannot play
      // this engine (src/xboard.c:106-117)
      if (strstr(line, "usermove") && !strcmp(cmd, "rejected")) {
        out("Error (unsupported feature usermove): rejected\n");
        out("tellusererror (unsupported feature usermove): rejected\n");
        quitengine(EXIT_FAILURE);
      }
    } else if (!strcmp(cmd, "new")) {
      if (!cl_release_device() || !cl_init_device((char *)"alphabeta_gpu")) quitengine(EXIT_FAILURE);
      new_game(STARTFEN);
      SD = MAXPLY;
      MaxNodes = nodes_for(MaxTime);
      if (!xboard_mode) printboard(BOARD, STM, GAMEPLY, PLY);
      xboard_force = false;
    } else if (!strcmp(cmd, "setboard")) {
      const char *fen = line + strlen("setboard");
      while (*fen == ' ') fen++;
      new_game(fen);
      if (!xboard_mode) printboard(BOARD, STM, GAMEPLY, PLY);
    } else if (!strcmp(cmd, "go")) {
      // only CECP v2 and later (src/xboard.c:217-226)
      if (xboard_mode && xboard_protover < 2) {
        out("Error (unsupported protocoll version, < v2): go\n");
        out("tellusererror (unsupported protocoll version. < v2): go\n");
        continue;
      }
      xboard_force = false;
      think();
    } else if (!strcmp(cmd, "force")) {
      xboard_force = true;
    } else if (time_command(cmd, line)) {
    } else if (!strcmp(cmd, "otim") || !strcmp(cmd, "memory")) {
    } else if (!strcmp(cmd, "usermove")) {
      // told, but the move is still played (src/xboard.c:464-476)
      if (xboard_mode && xboard_protover < 2) {
        out("Error (unsupported protocoll version, < v2): usermove\n");
        out("tellusererror (unsupported protocoll version, <v2): usermove\n");
      }
      char mv[16] = {0};
      sscanf(line, "usermove %15s", mv);
      const uint32_t m = can2move(mv, BOARD, STM);
      if (!m) {
        out("Illegal move: %s\n", mv);
        continue;
      }
      push_move(m);
      if (!xboard_mode || xboard_debug) printboard(BOARD, STM, GAMEPLY, PLY);
      if (!xboard_force) think();
    } else if (!strcmp(cmd, "undo")) {
      if (PLY > 0) pop_move();
    } else if (!strcmp(cmd, "remove")) {
      if (PLY >= 2) {
        pop_move();
        pop_move();
      }
    } else if (!strcmp(cmd, "quit")) {
      break;
    } else if (!strcmp(cmd, "sd")) {
      sscanf(line, "sd %d", &SD);
      SD = SD > MAXPLY ? MAXPLY : SD;
    } else if (!strcmp(cmd, "post")) {
      xboard_post = true;
    } else if (!strcmp(cmd, "nopost")) {
      xboard_post = false;
    } else if (!strcmp(cmd, "ping")) {
      uint64_t pong = 0;
      sscanf(line, "ping %" SCNu64, &pong);
      out("pong %" PRIu64 "\n", pong);
    } else if (!strcmp(cmd, "random") || !strcmp(cmd, "white") || !strcmp(cmd, "black") || !strcmp(cmd, "draw") ||
               !strcmp(cmd, "result") || !strcmp(cmd, "hint") || !strcmp(cmd, "bk") || !strcmp(cmd, "hard") ||
               !strcmp(cmd, "easy") || !strcmp(cmd, "name") || !strcmp(cmd, "rating") || !strcmp(cmd, "ics") ||
               !strcmp(cmd, "computer")) {
    } else if (!xboard_mode && !strcmp(cmd, "perft")) {
      out("### doing inits for perft depth %d: ###\n", SD);
      if (!cl_release_device() || !cl_init_device((char *)"perft_gpu")) quitengine(EXIT_FAILURE);
      out("### computing perft depth %d: ###\n", SD);
      const double start = now_ms();
      perft(BOARD, STM, SD);
      const double el = (now_ms() - start + 1) / 1000;
      if (!cl_release_device()) quitengine(EXIT_FAILURE);
      out("nodecount:%" PRIu64 ", seconds: %lf, nps: %" PRIu64 " \n", ABNODECOUNT, el,
          (uint64_t)((double)ABNODECOUNT / el));
      if (!cl_init_device((char *)"alphabeta_gpu")) quitengine(EXIT_FAILURE);
    } else if (!xboard_mode && !strcmp(cmd, "benchsmp")) {
      // the current position with 1, 2, 4, ... workers (src/xboard.c:822-974)
      std::vector<uint64_t> ws;
      smp_runs("benchsmp", ws);
      if (!cl_release_device() || !cl_init_device((char *)"alphabeta_gpu")) quitengine(EXIT_FAILURE);
    } else if (!xboard_mode && !strcmp(cmd, "benchkaufmann")) {
      bench_suite("benchkaufmann", KAUFMANN, 25);
    } else if (!xboard_mode && !strcmp(cmd, "benchhyatt24")) {
      bench_suite("benchhyatt24", HYATT24, 24);
    } else if (!xboard_mode && !strcmp(cmd, "bench")) {
      // fixed-depth search of the current position (time to depth and nps)
      const double start = now_ms();
      rootsearch(BOARD, STM, SD);
      const double el = (now_ms() - start) / 1000;
      out("#> bench: depth %d, %" PRIu64 " nodes in %lf seconds, nps: %" PRIu64 "\n", SD, ABNODECOUNT, el,
          (uint64_t)(ABNODECOUNT / (el > 0 ? el : 1e-9)));
    } else if (!xboard_mode && !strcmp(cmd, "selftest")) {
      if (!cl_release_device() || !cl_init_device((char *)"perft_gpu")) quitengine(EXIT_FAILURE);
      selftest();
      if (!cl_release_device() || !cl_init_device((char *)"alphabeta_gpu")) quitengine(EXIT_FAILURE);
    } else if (!xboard_mode && !strcmp(cmd, "help")) {
      print_help();
    } else if (!xboard_mode && !strcmp(cmd, "log")) {
      if (!LogFile) LogFile = fopen(LOGFILE, "a");
      else {
        fclose(LogFile);
        LogFile = nullptr;
      }
    } else if (!strcmp(cmd, "analyze") || !strcmp(cmd, "pause") || !strcmp(cmd, "resume") ||
               !strcmp(cmd, "edit")) {
      // named but unsupported commands also tell the user (src/xboard.c:1464-1490)
      out("Error (unsupported command): %s\n", cmd);
      out("tellusererror (unsupported command): %s\n", cmd);
      if (!strcmp(cmd, "edit")) out("tellusererror engine supports only CECP (Xboard) version >=2\n");
    } else {
      out("Error (unsupported command): %s\n", cmd);
    }
  }
}

}  // namespace

int main(int argc, char **argv) {
  setvbuf(stdout, nullptr, _IONBF, 0);
  for (int c = 1; c < argc; c++)
    if (!strcmp(argv[c], "-l") || !strcmp(argv[c], "--log")) LogFile = fopen(LOGFILE, "a");
  // the reference's options (src/zeta.c:529-541), read as it reads them with
  // getopt_long_only (one or two dashes, unique prefixes: -h, -selftest,
  // --guessconfigx ...), plus --config FILE; an unknown option prints the help
  static const struct option opts[] = {
      {"help", no_argument, nullptr, 0},        {"version", no_argument, nullptr, 0},
      {"selftest", no_argument, nullptr, 0},    {"log", no_argument, nullptr, 0},
      {"guessconfig", no_argument, nullptr, 0}, {"guessconfigx", no_argument, nullptr, 0},
      {"p", required_argument, nullptr, 0},     {"d", required_argument, nullptr, 0},
      {"pl", no_argument, nullptr, 0},          {"dl", no_argument, nullptr, 0},
      {"config", required_argument, nullptr, 0}, {nullptr, 0, nullptr, 0}};
  bool want_selftest = false;
  for (int idx = 0, c; (c = getopt_long_only(argc, argv, "", opts, &idx)) != -1;) {
    if (c != 0) {
      print_help();
      return EXIT_FAILURE;
    }
    switch (idx) {
      case 0: print_help(); return EXIT_SUCCESS;
      case 1: out("Zeta version: %s\n", VERSION); return EXIT_SUCCESS;
      case 2: want_selftest = true; break;
      case 3: break;  // the log is opened above, before anything is printed
      case 4:
      case 5: return guess_config(idx == 5) ? EXIT_SUCCESS : EXIT_FAILURE;
      case 6: opencl_platform_id = atoi(optarg); break;
      case 7: opencl_device_id = atoi(optarg); break;
      case 8:
      case 9:
        // src/clquery.c lists OpenCL platforms and devices; this build has HIP
        // devices, the one -d names for guessconfig
        out("#> no OpenCL platforms or devices: this build runs on HIP (device %d, -d N)\n", opencl_device_id);
        return EXIT_SUCCESS;
      case 10: configfile = optarg; break;
      default: break;
    }
  }
  out("#> Zeta %s\n#> Chess engine, HIP build for MI355X.\n#> engine is initialising...\n", VERSION);
  out("feature done=0\n");
  if (!read_and_init_config(configfile)) quitengine(EXIT_FAILURE);
  if (!engineinits()) quitengine(EXIT_FAILURE);
  new_game(STARTFEN);
  if (want_selftest) {
    if (!cl_init_device((char *)"perft_gpu")) quitengine(EXIT_FAILURE);
    quitengine(selftest() ? EXIT_SUCCESS : EXIT_FAILURE);
  }
  if (!cl_init_device((char *)"alphabeta_gpu")) quitengine(EXIT_FAILURE);
  memcpy(GLOBAL_BOARD, BOARD, sizeof BOARD);
  if (!cl_write_objects() || !cl_run_alphabeta(STM, 0, 1)) quitengine(EXIT_FAILURE);
  out("#> ...finished basic inits.\n");
  xboard();
  quitengine(EXIT_SUCCESS);
}
