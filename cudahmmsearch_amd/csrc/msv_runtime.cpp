// msv_runtime.cpp -- host runtime behind include/msv_gpu.h.
//
// Owns device memory, the resident length-sorted target database, per-profile
// device tables, the launch schedule (plain or batched/persistent), hipGraph
// capture of batched searches, and the host tables that turn the device's
// integer results into usc and the F1 pass bit with the reference's own float
// and double expressions:
//   usc     impl_sse/msvfilter.c:208-210
//   nullsc  p7_bg.c:191-198 (p7_bg_SetLength) and :359-363 (p7_bg_NullOne)
//   pass    p7_pipeline.c:667-677 with esl_gumbel_surv, easel/esl_gumbel.c:129-137
//   tjb_b   impl_sse/p7_oprofile.c:921-925 via unbiased_byteify :512-519
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <limits>
#include <array>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/msv_gpu.h"
#include "msv_device.h"
#include "msv_host.h"
#include "../../include/msv_prep.h"

// Scratch memory of the allocation-free search (msvg_search_noalloc): one
// array per GPU, held until the process ends.  Host code never uses this
// name as an address: msvg_init looks up each device's copy with
// hipGetSymbolAddress and keeps that pointer in the context.
constexpr size_t kScratchBytes = (size_t) 128 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

extern "C" hipError_t msvk_prepare();
extern "C" hipError_t msvk_prepare_batch();
extern "C" hipError_t msvk_prepare_scan_group();
extern "C" int msvk_scan_group_built(int G, int QB);
extern "C" hipError_t msvk_launch_scan_group(int G, int QB, const ScanGroupArgs *a, int grid, int block, size_t lds,
                                             hipStream_t st);
extern "C" hipError_t msvk_launch_batch(int G, const BatArgs *a, int grid, int block, size_t lds, hipStream_t st);
extern "C" hipError_t msvk_batch_occupancy(int G, int block, size_t lds, int *per_cu);
extern "C" hipError_t msvk_build_bat_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int G, int Q,
                                           int bias, uint32_t *out, hipStream_t st);
extern "C" hipError_t msvk_launch_fast(int G, int Q, const FastArgs *a, int grid, int block, size_t lds, hipStream_t st);
extern "C" hipError_t msvk_launch_seg(int Q, const SegArgs *a, int grid, int block, size_t lds, hipStream_t st);
extern "C" hipError_t msvk_launch_vit_seg(int Q, const VitSegArgs *a, int grid, hipStream_t st);
extern "C" int msvk_vit_seg_cap(int Q, int R, int ncu);
extern "C" hipError_t msvk_launch_slow(const SlowArgs *a, int grid, hipStream_t st);
extern "C" hipError_t msvk_launch_jrow(int Q, const JArgs *a, int grid, size_t lds, hipStream_t st);
extern "C" hipError_t msvk_launch_jrow_many(const JManyArgs *a, size_t lds, hipStream_t st);
extern "C" hipError_t msvk_jrow_many_occupancy(size_t lds, int *per_cu);
extern "C" hipError_t msvk_launch_vit(int Q, int W, const VitArgs *a, int ncu, hipStream_t st);
extern "C" int msvk_vit_slots(int Q, int W, const VitArgs *a, int ncu);
extern "C" hipError_t msvk_launch_vit_band(int G, int Q, int vote, const VitArgs *a, int grid, hipStream_t st);
extern "C" int msvk_vit_band_block(int G);
extern "C" int msvk_vit_band_cap(int G, int Q, int vote, int R, int ncu);
extern "C" hipError_t msvk_launch_finalize(const FinArgs *a, int np, int64_t n, hipStream_t st);
extern "C" hipError_t msvk_build_fast_table(const uint8_t *rbv, int M, int Kp, const uint8_t *rowcode, int R, int G,
                                            int Q, int W, int bias, uint32_t *out, hipStream_t st);
extern "C" hipError_t msvk_build_seg_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int Q,
                                           int k0, int bias, uint32_t *out, hipStream_t st);
extern "C" hipError_t msvk_build_slow_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int Qs,
                                            uint32_t *out, hipStream_t st);
extern "C" hipError_t msvk_fill_zero(int32_t *p, int n, hipStream_t st);
extern "C" hipError_t msvk_lds_probe(void *out, int grid, int iters, hipStream_t st);
extern "C" hipError_t msvk_forms_probe(int form, const uint8_t *res, int L, const void *tab, int tab_u4,
                                       uint32_t bias2, uint32_t *out, int grid, hipStream_t st);
extern "C" hipError_t msvk_launch_one(int Q, const OneArgs *a, size_t lds, hipStream_t st);
extern "C" hipError_t msvk_launch_scan(int Q, const ScanArgs *a, int grid, size_t lds, hipStream_t st);

namespace {

constexpr int kLdsBytes = 160 * 1024;
constexpr int kQMax = 32;           // registers per lane in the fast kernels
constexpr int kVitStreams = 4;      // streams of a multi-profile ViterbiFilter search (the default hardware queues)
constexpr int kJStreams = 4;        // streams of a batched search's J-aware launches
constexpr int kSegQMax = 28;        // registers per lane in one model segment (at most 3584 positions)

// Segment width: the most registers whose table fits in LDS for R rows.
int seg_q(int R) {
  const int qb = kLdsBytes / ((R + 1) * 1024);
  return std::max(1, std::min(kSegQMax, 4 * qb));
}

// Failure injection for tests: MSVG_TEST_FAIL_ALLOC=k makes every device
// allocation after the first k fail as if the device were out of memory
// (checked on the host; no device call is made for the failing request).
std::atomic<long> g_alloc_budget{-2};

bool alloc_allowed() {
  long b = g_alloc_budget.load();
  if (b == -2) {
    const char *env = std::getenv("MSVG_TEST_FAIL_ALLOC");
    b = env ? std::atol(env) : -1;
    long expect = -2;
    g_alloc_budget.compare_exchange_strong(expect, b);
    b = g_alloc_budget.load();
  }
  if (b < 0) return true;
  return g_alloc_budget.fetch_sub(1) > 0;
}

// A bump allocator over a fixed device range (the allocation-free search
// carves its tables, queues and result words from the library's static
// scratch array): nothing is freed, a call starts again from the bottom.
struct Arena {
  uint8_t *base = nullptr;
  size_t cap = 0, used = 0;
  void *take(size_t bytes) {
    const size_t off = (used + 255) & ~(size_t) 255;
    if (off > cap || bytes > cap - off) return nullptr;
    used = off + bytes;
    return base + off;
  }
};

struct DevMem {
  void *p = nullptr;
  size_t n = 0;
  Arena *arena = nullptr;          // non-null: memory comes from this arena (never freed here)
  bool borrowed = false;           // p is the caller's memory
  DevMem() = default;
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem() { if (p && !arena && !borrowed) (void) hipFree(p); }
  hipError_t ensure(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    if (arena || borrowed) {       // no allocation: the range given must do
      if (p || borrowed) return hipErrorOutOfMemory;
      p = arena->take(bytes ? bytes : 16);
      if (!p) return hipErrorOutOfMemory;
      n = bytes;
      return hipSuccess;
    }
    if (p) { (void) hipFree(p); p = nullptr; n = 0; }
    if (!alloc_allowed()) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e == hipSuccess) n = bytes;
    return e;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct VitProfile {               // ViterbiFilter fields (model order)
  std::vector<int16_t> rwv, twv;   // Kp * M, 8 * M
  int e_loop = 0, e_move = 0, base_w = 0, ddbound_w = 0;
  float scale_w = 0.f, nj = 0.f, vmu = 0.f, vlambda = 0.f;
  int Q = 1, W = 1;                // register blocks per lane, waves per target
  int G = 64;                      // lanes per target: 16 or 32 (several targets per wave), or 64 * W
                                   // (W = 1: vit_band; W > 1: vit_kernel)
  int vote = 1;                    // vit_band: rows vote on the lazy-F test (else always close)
  DevMem d_ttab;                   // transition table, VitArgs layout
  // Multi-wave shapes (W > 1) also have column segments for vit_seg:
  // positions per segment (128 x at most 16 registers) and each segment's
  // transition table in the 64-lane layout.
  int seg_span = 0;
  std::vector<std::unique_ptr<DevMem>> d_seg_ttab;
};

struct Profile {
  int M = 0, Kp = 0;
  std::vector<uint8_t> rbv;        // Kp * M
  int tbm = 0, tec = 0, base = 0, bias = 0;
  float scale = 0.f, mu = 0.f, lambda = 0.f;
  DevMem d_rbv;
  std::unique_ptr<VitProfile> vit;
  DevMem d_onetab;                 // per-call drop-in: 64-lane fast table over codes 0..Kp-1
  DevMem d_scantab;                // scan direction: the same table at the profile's register class
  int scanQ = 0;                   // that class (0: not built)
  DevMem d_gtab;                   // batched scan: fast table in the G-lane layout over codes 0..Kp-1
  int gG = 0, gQ = 0;              // its lanes per target and registers (0: not built)
};

struct ScanSet {                   // the profiles of msvg_scan calls, grouped by register class
  std::vector<int32_t> pids;
  std::vector<Profile *> prof;     // per position (valid while the set lives: removal resets it)
  int Kp = 0;
  struct Cls {
    int Q = 0;
    std::vector<int32_t> members;  // positions in pids
    DevMem d_profs;                // ScanProf per member
    std::vector<ScanProf> h_profs;
    std::vector<int32_t> slot;     // position in pids -> member index, -1 if not a member
  };
  std::vector<std::unique_ptr<Cls>> cls;
  // Per class member, in class order: its position and the constants the
  // host's decoding reads (one sequential array rather than each profile's
  // record); rebuilt after msvg_profile_set_evparams changes mu or lambda.
  struct Row { int32_t i; int base; float scale, mu, lambda; };
  std::vector<Row> rows;
  bool rows_ok = false;
  std::vector<Row> grows;          // the same in the grouped classes' order (gcls)
  bool grows_ok = false;
  std::vector<int32_t> fallback;   // positions whose table does not fit LDS
  int64_t sumM = 0;
  // Batched scans: the same profiles grouped by G lanes per query and QB
  // register blocks (msv_scan_group<G, QB>), members by register count,
  // most first; built on the first batched call.
  struct GCls {
    int G = 0, QB = 0;
    int Qmax = 0;                  // registers of the first member
    int64_t Qsum = 0;              // over the members
    std::vector<int32_t> members;
    DevMem d_profs;                // ProfDev per member
  };
  std::vector<std::unique_ptr<GCls>> gcls;
  bool grouped = false;
};

struct Db {
  int64_t n = 0, total_res = 0, packed = 0;
  int R = 0;
  std::vector<uint8_t> rowcode;         // row -> residue code
  int maxcode = -1;
  std::vector<int32_t> len_sorted;
  std::vector<int64_t> sorted_of;       // caller index -> sorted index
  std::vector<int32_t> Ls;              // distinct lengths
  std::vector<int32_t> lidx_sorted;
  std::vector<uint8_t> tjb_sorted;
  DevMem d_res, d_off, d_len, d_tjb, d_lidx, d_sorted_pos, d_order, d_rowcode;
  std::map<int, std::pair<std::unique_ptr<DevMem>, int>> chunks;   // G -> (chunk_start, nchunks)
};

struct Tables {                          // per (profile, db)
  std::map<int, std::unique_ptr<DevMem>> fast;   // G -> table
  std::map<int, std::unique_ptr<DevMem>> bat;    // G -> batched-pass table (G = 8, 16, 32)
  std::vector<std::unique_ptr<DevMem>> seg;      // segment tables (M > fast limit)
  std::unique_ptr<DevMem> slow;
  std::unique_ptr<DevMem> vemit;                 // ViterbiFilter emission table
  std::vector<std::unique_ptr<DevMem>> vseg_emit; // ViterbiFilter column segments' emission tables
};

struct Tile { int lo, hi, lanes, waves; };

struct Bufs {                             // device buffers one search writes
  DevMem d_res, d_prof, d_slow, d_counters, d_usc, d_status, d_pass, d_thr, d_usctab, d_seg, d_jslow;
  DevMem d_bat;                          // batched launches: BatProf arrays and first-profile lists
  DevMem d_jprof;                        // batched searches' J stage: JProf per profile
  int64_t slow_cap = 0;
  int64_t jcap = 0;                      // per-profile J-aware queue capacity
  Arena *arena = nullptr;                // allocation-free search: tables, chunk lists and buffers from here
  bool ext_out = false;                  // usc/status/pass are the caller's device arrays (any may be null)
  Bufs() = default;
  explicit Bufs(Arena *a) : arena(a) {
    for (DevMem *m : {&d_res, &d_prof, &d_slow, &d_counters, &d_usc, &d_status, &d_pass, &d_thr, &d_usctab,
                      &d_seg, &d_jslow, &d_bat, &d_jprof})
      m->arena = a;
  }
};

struct Plan {                             // a captured batched search with its own buffers
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  Bufs bufs;
  ~Plan() {
    if (exec) (void) hipGraphExecDestroy(exec);
    if (graph) (void) hipGraphDestroy(graph);
  }
};

}  // namespace

struct msvg_ctx {
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;      // batched mode: J-aware launches beside the fast chain
  hipEvent_t fork = nullptr, join = nullptr;
  hipStream_t jst[kJStreams] = {}; // batched mode: the J-aware launches after the batched pass
  hipEvent_t jdone[kJStreams] = {};
  std::string err;
  int next_id = 1;
  std::map<int, std::unique_ptr<Profile>> profiles;
  std::map<int, std::unique_ptr<Db>> dbs;
  std::map<std::pair<int, int>, std::unique_ptr<Tables>> tables;
  std::map<std::tuple<int, int, double>, std::pair<std::vector<uint16_t>, std::vector<float>>> finals;
  std::vector<Tile> tiles;
  std::vector<Tile> vtiles;        // ViterbiFilter shapes by model length (msvg_set_vit_tile)
  int vit_vote = -1;               // msvg_set_vit_vote: -1 by the profile's lazy-F rate, else forced
  int vit_segments = 1;            // msvg_set_vit_segments: multi-wave shapes run as column segments
  bool tail_guard = true;
  Bufs eager;                      // buffers of plain and eager searches
  Bufs *last = nullptr;            // buffers holding the latest results
  int64_t slow_cap = 0;            // slow-queue capacity for the next search
  int last_np = 0, last_db = -1;
  uint8_t *scratch_dev = nullptr;  // this device's copy of the scratch array (hipGetSymbolAddress)
  int last_launches = 0;           // J-free launches of the latest enqueued search
  int last_nbatch = 0;             // batched launches of the latest enqueued search
  int last_batch_G[4] = {}, last_batch_timed = 0;
  int64_t last_batch_cells[4] = {};
  hipEvent_t bev[10] = {};         // eager batched searches: around each batched launch, then the J stage
  msvg_stats stats{};
  hipEvent_t ev[8] = {};
  std::map<std::string, std::unique_ptr<Plan>> plans;
  // single-target drop-in
  int scratch_db = -1;
  DevMem d_one_res, d_one_out, d_idrow;     // target rows, result word, identity row map
  uint8_t *h_one_res = nullptr;             // pinned staging for the target
  size_t h_one_cap = 0;
  uint32_t *h_one_out = nullptr;            // pinned result word
  // scan direction (many profiles, one query)
  std::unique_ptr<ScanSet> scanset;
  DevMem d_scanout;
  DevMem d_scanj;                           // batched scan's J stage: profile lists, query lists, words
  uint32_t *h_scanout = nullptr;            // pinned result words
  size_t h_scancap = 0;
  // ViterbiFilter searches
  DevMem d_vitems, d_vmove, d_vout, d_vuoff, d_vprofs;
  DevMem d_vcoff, d_vpartial, d_vcarry;   // vit_seg: carry offsets, partial maxima, carries
  int64_t vcrows = 0;                       // rows of each of the four carry arrays in d_vcarry
  hipStream_t vstream[kVitStreams] = {};   // multi-profile searches: launches round-robin
  hipEvent_t vdone[kVitStreams] = {};
};

namespace {

int fail(msvg_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hipfail(msvg_ctx *c, hipError_t e, const char *what) {
  return fail(c, MSVG_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                   \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return hipfail((ctx), _e, #expr); \
  } while (0)

int qfor(int M, int slots) { return (M + slots - 1) / slots; }

// Lane-group size and waves per workgroup for a profile of length M.
void choose_tile_tp(const msvg_ctx *c, int M, int R, int &G, int &waves) {
  // An installed tile applies only where its table fits this database's
  // LDS footprint; otherwise the cost model below decides.
  for (const Tile &t : c->tiles)
    if (M >= t.lo && M <= t.hi) {
      const int Q = qfor(M, 2 * t.lanes);
      if (Q <= kQMax && (size_t) (R + 1) * ((Q + 3) / 4) * 1024 <= (size_t) kLdsBytes) {
        G = t.lanes;
        waves = t.waves;
        return;
      }
      break;
    }
  // Built-in cost model: VALU instructions per wave-row divided by the model
  // cells the wave advances, over the sizes whose tables fit in LDS.
  double best = 1e30;
  G = 64;
  for (int g : {8, 16, 32, 64}) {
    const int Q = qfor(M, 2 * g);
    if (Q > kQMax) continue;
    const size_t lds = (size_t) (R + 1) * ((Q + 3) / 4) * 1024;
    if (lds > (size_t) kLdsBytes) continue;
    const double v = Q + (Q + 1) / 2 + 4 + ((g == 32 || g == 8) ? 1 : 0);
    const double cost = v / ((64 / g) * (double) M);
    if (cost < best - 1e-12) { best = cost; G = g; }
  }
  waves = 0;
}

size_t fast_lds(int R, int Q) { return (size_t) (R + 1) * ((Q + 3) / 4) * 1024; }

// The tile for profile length M over database d.  choose_tile_tp picks for
// throughput; a launch also lasts at least as long as its longest target's
// row chain, which one lane group walks alone at the end.  Where that chain
// outlasts the throughput estimate (small databases holding a long target),
// each lane-group size is scored by max(chain, throughput time) and the
// best is taken if it beats the throughput tile by 15%.  Models, fitted on
// one MI355X: a row of a lane group costs v = Q + ceil(Q/2) packed
// instructions; the chain takes 2.3 ns per instruction + 33 ns per row; the
// device retires 430 wave-instructions per ns, and a target row costs
// (v + 4) * G / 64 of them.
void choose_tile(const msvg_ctx *c, int M, const Db &d, int &G, int &waves) {
  choose_tile_tp(c, M, d.R, G, waves);
  if (!c->tail_guard || d.len_sorted.empty()) return;
  const double Lmax = (double) ((d.len_sorted.front() + 15) & ~15);
  auto cost = [&](int g) {
    const int Q = qfor(M, 2 * g);
    const double v = Q + (Q + 1) / 2;
    const double chain = Lmax * (2.3 * v + 33.0);
    const double thru = (double) d.total_res * (v + 4.0) * g / 64.0 / 430.0;
    return std::make_pair(chain, thru);
  };
  const auto c0 = cost(G);
  if (c0.first <= c0.second) return;
  double best = std::max(c0.first, c0.second);
  int bestG = G;
  for (int g : {8, 16, 32, 64}) {
    const int Q = qfor(M, 2 * g);
    if (g == G || Q > kQMax || fast_lds(d.R, Q) > (size_t) kLdsBytes) continue;
    const auto cg = cost(g);
    const double t = std::max(cg.first, cg.second);
    if (t < best) { best = t; bestG = g; }
  }
  if (bestG != G && best < 0.85 * std::max(c0.first, c0.second)) {
    G = bestG;
    waves = 0;
  }
}

// Workgroup size and grid for a persistent launch whose table uses lds bytes.
void persistent_shape(const msvg_ctx *c, size_t lds, int waves_hint, int &grid, int &block) {
  int per_cu = (int) std::max<size_t>(1, std::min<size_t>(8, kLdsBytes / std::max<size_t>(lds, 1)));
  int waves = waves_hint > 0 ? waves_hint : std::min(16, std::max(4, (16 + per_cu - 1) / per_cu));
  block = 64 * waves;
  grid = c->ncu * per_cu;
}

float usc_of(int d, int base, float scale) {
  // msvfilter.c:208-210 with xJ - tjb_b = d
  float sc = ((float) d - (float) base);
  sc /= scale;
  sc -= 3.0;
  return sc;
}

// The F1 pass decision of p7_pipeline.c:667-677 for a finite usc.
bool passes(float usc, int L, float mu, float lambda, double F1) {
  const float nullsc = msvh_nullsc(L);
  const float seq_score = (usc - nullsc) / 0.69314718055994529;
  const double y = (double) lambda * ((double) seq_score - (double) mu);
  const double ey = -std::exp(-y);
  const double P = (std::fabs(ey) < 5e-9) ? -ey : 1 - std::exp(ey);
  return !(P > F1);
}

// passes() for many scores of one target length: nullsc(L) computed once,
// and the Gumbel survival evaluated only near its F1 crossing.  P falls
// strictly with y = lambda (seq_score - mu), so the rule !(P > F1) is y >= y*
// with y* = -log(-log(1 - F1)); scores more than 1e-6 from y* are decided by
// the comparison (P's double evaluation is accurate to ~1e-16 there), the
// rest by the exact expression.
struct PassRule {
  float nullsc;
  double F1, ystar;
  bool fast;
  PassRule(int L, double f1) : nullsc(msvh_nullsc(L)), F1(f1), ystar(0), fast(f1 > 0 && f1 < 1) {
    if (fast) ystar = -std::log(-std::log1p(-f1));
  }
  bool operator()(float usc, float mu, float lambda) const {
    const float seq_score = (usc - nullsc) / 0.69314718055994529;
    const double y = (double) lambda * ((double) seq_score - (double) mu);
    if (fast && y > ystar + 1e-6) return true;
    if (fast && y < ystar - 1e-6) return false;
    const double ey = -std::exp(-y);
    const double P = (std::fabs(ey) < 5e-9) ? -ey : 1 - std::exp(ey);
    return !(P > F1);
  }
};

const std::pair<std::vector<uint16_t>, std::vector<float>> &final_tables(msvg_ctx *c, int pid, int dbid, double F1) {
  auto key = std::make_tuple(pid, dbid, F1);
  auto it = c->finals.find(key);
  if (it != c->finals.end()) return it->second;
  const Profile &p = *c->profiles[pid];
  const Db &d = *c->dbs[dbid];
  std::vector<float> usc(511);
  for (int k = 0; k < 511; ++k) usc[k] = usc_of(k - 255, p.base, p.scale);
  std::vector<uint16_t> thr(d.Ls.size());
  int guess = 128;
  for (size_t li = 0; li < d.Ls.size(); ++li) {
    const int L = d.Ls[li];
    const int tjb = msvh_tjb(L);
    auto ok = [&](int xJ) { return passes(usc[xJ - tjb + 255], L, p.mu, p.lambda, F1); };
    int t;
    if (!ok(255)) t = 256;
    else if (ok(0)) t = 0;
    else {
      // smallest passing xJ: the score is monotone in xJ.  Start from the
      // previous length's threshold and widen, then bisect.
      int lo = std::max(0, guess - 2), hi = std::min(255, guess + 2);
      while (lo > 0 && ok(lo)) lo = std::max(0, lo - 8);
      while (!ok(hi)) hi = std::min(255, hi + 8);
      if (ok(lo)) t = lo;  // lo == 0 passing handled above; keep for safety
      else {
        while (hi - lo > 1) { int m = (lo + hi) / 2; if (ok(m)) hi = m; else lo = m; }
        t = hi;
      }
    }
    thr[li] = (uint16_t) t;
    if (t <= 255) guess = t;
  }
  return c->finals.emplace(key, std::make_pair(std::move(thr), std::move(usc))).first->second;
}

std::vector<int32_t> guided_chunks(const msvg_ctx *c, const Db &d, int G) {
  const int NG = 64 / G;
  const int64_t ntasks = (d.n + NG - 1) / NG;
  // Guided chunks: each holds about 1/(8 waves) of the residues not yet
  // chunked (a full chip runs ncu * 16 waves), down to 2,048 residues, so
  // chunk sizes shrink toward the end of the (longest-first) task list and
  // the waves finish within a small chunk of each other.
  const int64_t waves = (int64_t) c->ncu * 16;
  int64_t remaining = 0;
  for (int64_t t = 0; t < ntasks; ++t) remaining += d.len_sorted[t * NG];
  std::vector<int32_t> cs;
  cs.push_back(0);
  int64_t acc = 0;
  int64_t target = std::max<int64_t>(2048, remaining / (waves * 8));
  for (int64_t t = 0; t < ntasks; ++t) {
    acc += d.len_sorted[t * NG];     // the longest in the task sets its rows
    if (acc >= target) {
      cs.push_back((int32_t) (t + 1));
      remaining -= acc;
      acc = 0;
      target = std::max<int64_t>(2048, remaining / (waves * 8));
    }
  }
  if (cs.back() != ntasks) cs.push_back((int32_t) ntasks);
  return cs;
}

int build_chunks(msvg_ctx *c, Db &d, int G) {
  auto it = d.chunks.find(G);
  if (it != d.chunks.end()) return MSVG_OK;
  const std::vector<int32_t> cs = guided_chunks(c, d, G);
  auto mem = std::make_unique<DevMem>();
  HIPCHK(c, mem->ensure(cs.size() * sizeof(int32_t)));
  HIPCHK(c, hipMemcpy(mem->p, cs.data(), cs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  d.chunks.emplace(G, std::make_pair(std::move(mem), (int) cs.size() - 1));
  return MSVG_OK;
}

// The chunk list of lane-group size G over database d: cached on the
// database, or (allocation-free search) copied into the arena per call.
int chunk_list(msvg_ctx *c, Db &d, int G, Arena *ar, const int32_t **cs_out, int *nchunks) {
  if (!ar) {
    int rc = build_chunks(c, d, G);
    if (rc != MSVG_OK) return rc;
    *cs_out = d.chunks[G].first->as<int32_t>();
    *nchunks = d.chunks[G].second;
    return MSVG_OK;
  }
  const std::vector<int32_t> cs = guided_chunks(c, d, G);
  void *m = ar->take(cs.size() * sizeof(int32_t));
  if (!m) return fail(c, MSVG_EINVAL, "allocation-free search: scratch exhausted (chunk list)");
  HIPCHK(c, hipMemcpy(m, cs.data(), cs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  *cs_out = static_cast<const int32_t *>(m);
  *nchunks = (int) cs.size() - 1;
  return MSVG_OK;
}

Tables &tables_for(msvg_ctx *c, int pid, int dbid) {
  auto &t = c->tables[{pid, dbid}];
  if (!t) t = std::make_unique<Tables>();
  return *t;
}

int ensure_fast_table(msvg_ctx *c, int pid, int dbid, int G, const uint32_t **out, hipStream_t st,
                      Arena *ar = nullptr) {
  Profile &p = *c->profiles[pid];
  Db &d = *c->dbs[dbid];
  if (ar) {                                 // allocation-free search: built per call in the arena
    const int Q = qfor(p.M, 2 * G);
    void *m = ar->take(fast_lds(d.R, Q));
    if (!m) return fail(c, MSVG_EINVAL, "allocation-free search: scratch exhausted (fast table)");
    HIPCHK(c, msvk_build_fast_table(p.d_rbv.as<uint8_t>(), p.M, p.Kp, d.d_rowcode.as<uint8_t>(), d.R, G, Q, 64, p.bias,
                                    static_cast<uint32_t *>(m), st));
    *out = static_cast<const uint32_t *>(m);
    return MSVG_OK;
  }
  Tables &t = tables_for(c, pid, dbid);
  auto it = t.fast.find(G);
  if (it == t.fast.end()) {
    const int Q = qfor(p.M, 2 * G);
    auto mem = std::make_unique<DevMem>();
    HIPCHK(c, mem->ensure(fast_lds(d.R, Q)));
    HIPCHK(c, msvk_build_fast_table(p.d_rbv.as<uint8_t>(), p.M, p.Kp, d.d_rowcode.as<uint8_t>(), d.R, G, Q, 64, p.bias,
                                    mem->as<uint32_t>(), st));
    it = t.fast.emplace(G, std::move(mem)).first;
  }
  *out = it->second->as<uint32_t>();
  return MSVG_OK;
}

// Batched-pass tables (msv_device.h, BatProf): G = 64 uses the fast-stage
// table; G = 8, 16, 32 the narrow [block][32 rows][TW lanes] layout.
int bat_tw(int G) { return G < 16 ? 16 : G; }
// Byte-MSV queue of the allocation-free search (fixed: overflow is an error).
int64_t noalloc_slow_cap(int np, int64_t n) { return std::max<int64_t>(4096, (int64_t) np * n / 16); }
int bat_qmax(int G) { return G == 64 ? kQMax : G == 16 ? 24 : 16; }
size_t bat_bytes(int G, int R, int Q) {
  return G == 64 ? fast_lds(R, Q) : (size_t) ((Q + 3) / 4) * 32 * bat_tw(G) * 16;
}
// Does the batched kernel take this profile at lane-group size G?  (Its
// table plus the workgroup's broadcast word fit LDS; narrow layouts need
// at most 32 residue rows and bat_qmax registers.)
bool bat_fits(int G, int M, int R) {
  const int Q = qfor(M, 2 * G);
  if (Q > bat_qmax(G)) return false;
  if (G != 64 && R + 1 > 32) return false;
  return bat_bytes(G, R, Q) + 16 <= (size_t) kLdsBytes;
}

// Lane-group size of a profile of length M in a batched search over d: the
// tile's G, or the next wider one the batched kernel takes; 0 when none does
// (column-segmented models, tables at the LDS limit).
int batch_lanes(const msvg_ctx *c, int M, const Db &d) {
  int G = 64, w = 0;
  choose_tile(c, M, d, G, w);
  const int Q = qfor(M, 2 * G);
  if (Q > kQMax || fast_lds(d.R, Q) > (size_t) kLdsBytes) return 0;
  while (G < 64 && !bat_fits(G, M, d.R)) G *= 2;
  return bat_fits(G, M, d.R) ? G : 0;
}

int ensure_bat_table(msvg_ctx *c, int pid, int dbid, int G, const uint32_t **out, hipStream_t st,
                     Arena *ar = nullptr) {
  if (G == 64) return ensure_fast_table(c, pid, dbid, 64, out, st, ar);
  Profile &p = *c->profiles[pid];
  Db &d = *c->dbs[dbid];
  if (ar) {
    const int Q = qfor(p.M, 2 * G);
    void *m = ar->take(bat_bytes(G, d.R, Q));
    if (!m) return fail(c, MSVG_EINVAL, "allocation-free search: scratch exhausted (batched table)");
    HIPCHK(c, msvk_build_bat_table(p.d_rbv.as<uint8_t>(), p.M, d.d_rowcode.as<uint8_t>(), d.R, G, Q, p.bias,
                                   static_cast<uint32_t *>(m), st));
    *out = static_cast<const uint32_t *>(m);
    return MSVG_OK;
  }
  Tables &t = tables_for(c, pid, dbid);
  auto it = t.bat.find(G);
  if (it == t.bat.end()) {
    const int Q = qfor(p.M, 2 * G);
    auto mem = std::make_unique<DevMem>();
    HIPCHK(c, mem->ensure(bat_bytes(G, d.R, Q)));
    HIPCHK(c, msvk_build_bat_table(p.d_rbv.as<uint8_t>(), p.M, d.d_rowcode.as<uint8_t>(), d.R, G, Q, p.bias,
                                   mem->as<uint32_t>(), st));
    it = t.bat.emplace(G, std::move(mem)).first;
  }
  *out = it->second->as<uint32_t>();
  return MSVG_OK;
}

int ensure_seg_tables(msvg_ctx *c, int pid, int dbid, hipStream_t st) {
  Profile &p = *c->profiles[pid];
  Db &d = *c->dbs[dbid];
  Tables &t = tables_for(c, pid, dbid);
  if (!t.seg.empty()) return MSVG_OK;
  const int span = 128 * seg_q(d.R);
  for (int k0 = 0; k0 < p.M; k0 += span) {
    const int Mseg = std::min(span, p.M - k0);
    const int Q = qfor(Mseg, 128);
    auto mem = std::make_unique<DevMem>();
    HIPCHK(c, mem->ensure(fast_lds(d.R, Q)));
    HIPCHK(c, msvk_build_seg_table(p.d_rbv.as<uint8_t>(), p.M, d.d_rowcode.as<uint8_t>(), d.R, Q, k0, p.bias,
                                   mem->as<uint32_t>(), st));
    t.seg.push_back(std::move(mem));
  }
  return MSVG_OK;
}

int ensure_slow_table(msvg_ctx *c, int pid, int dbid, const uint32_t **out, hipStream_t st, Arena *ar = nullptr) {
  Profile &p = *c->profiles[pid];
  Db &d = *c->dbs[dbid];
  if (ar) {
    const int Qs = qfor(p.M, 128);
    void *m = ar->take((size_t) (d.R + 1) * Qs * 64 * 4);
    if (!m) return fail(c, MSVG_EINVAL, "allocation-free search: scratch exhausted (byte-stage table)");
    HIPCHK(c, msvk_build_slow_table(p.d_rbv.as<uint8_t>(), p.M, d.d_rowcode.as<uint8_t>(), d.R, Qs,
                                    static_cast<uint32_t *>(m), st));
    *out = static_cast<const uint32_t *>(m);
    return MSVG_OK;
  }
  Tables &t = tables_for(c, pid, dbid);
  if (!t.slow) {
    const int Qs = qfor(p.M, 128);
    auto mem = std::make_unique<DevMem>();
    HIPCHK(c, mem->ensure((size_t) (d.R + 1) * Qs * 64 * 4));
    HIPCHK(c, msvk_build_slow_table(p.d_rbv.as<uint8_t>(), p.M, d.d_rowcode.as<uint8_t>(), d.R, Qs,
                                    mem->as<uint32_t>(), st));
    t.slow = std::move(mem);
  }
  *out = t.slow->as<uint32_t>();
  return MSVG_OK;
}

int check_ids(msvg_ctx *c, const int32_t *ids, int np, int dbid) {
  if (!c) return MSVG_EINVAL;
  if (np < 1 || !ids) return fail(c, MSVG_EINVAL, "no profiles given");
  auto dit = c->dbs.find(dbid);
  if (dit == c->dbs.end()) return fail(c, MSVG_EINVAL, "unknown database id");
  for (int i = 0; i < np; ++i) {
    auto pit = c->profiles.find(ids[i]);
    if (pit == c->profiles.end()) return fail(c, MSVG_EINVAL, "unknown profile id");
    if (dit->second->maxcode >= pit->second->Kp)
      return fail(c, MSVG_EINVAL, "database holds residue codes outside the profile's alphabet");
  }
  return MSVG_OK;
}

// Enqueue one complete search on c->stream.  When timed, per-stage events are
// recorded (never inside a graph capture).
// When capturing, nothing may be copied from host memory (the graph would
// replay a copy of a buffer that no longer exists) and nothing may be
// synchronized; the warm-up call made just before the capture leaves every
// table and parameter block in b, and replays of the graph refill them
// first (Plan upload in search_device).
// sched: 0 persistent (batched mode), 2 one chunk per workgroup (serving:
// concurrent searches on separate streams share the CUs).
// side: when given (batched searches on c->stream), each profile's J-aware
// launch goes to this stream as soon as the profile's fast launch is done, so
// it runs beside the next profiles' fast launches instead of after them
// (each J-aware launch lasts as long as its longest pair, one wave per pair);
// the stream joins back before the byte stage.
int enqueue_search(msvg_ctx *c, Bufs &b, const int32_t *ids, int np, int dbid, double F1, int mode, bool timed,
                   bool capturing, int64_t *cells_out, hipStream_t st = nullptr, int sched = 0,
                   hipStream_t side = nullptr, bool tables_only = false) {
  if (!st) st = c->stream;
  bool side_used = false;
  // J-aware launch of profile i on the side stream (or in line on st).
  auto launch_j = [&](const JArgs &ja, int Q64, size_t jl, int grid) -> int {
    hipStream_t js = st;
    if (side) {
      HIPCHK(c, hipEventRecord(c->fork, st));
      HIPCHK(c, hipStreamWaitEvent(side, c->fork, 0));
      js = side;
      side_used = true;
    }
    HIPCHK(c, msvk_launch_jrow(Q64, &ja, grid, jl, js));
    return MSVG_OK;
  };
  Db &d = *c->dbs[dbid];
  const int64_t n = d.n;
  int rc;
  std::vector<ProfDev> pds(np);
  std::vector<int> Gs(np), waves(np), segmented(np);
  std::vector<int> Gb(np, 0);          // lane-group size of the profile's batched launch (0: none)
  std::vector<const uint4 *> jtab(np);
  // Batched searches on persistent launches take every profile the batched
  // kernel fits (msv_batch.hip); shared-schedule serving keeps msv_fast.
  const bool batch = (mode == MSVG_MODE_BATCHED && sched == 0);
  c->last_nbatch = 0;
  c->last_batch_timed = 0;
  int64_t cells = 0;
  for (int i = 0; i < np; ++i) {
    Profile &p = *c->profiles[ids[i]];
    ProfDev &pd = pds[i];
    int G = 64, w = 0;
    if (mode == MSVG_MODE_BATCHED) choose_tile(c, p.M, d, G, w);
    const int Q = qfor(p.M, 2 * G);
    segmented[i] = (Q > kQMax || fast_lds(d.R, Q) > (size_t) kLdsBytes) ? 1 : 0;
    if (segmented[i]) G = 64;
    if (!segmented[i] && batch) {
      const int g = batch_lanes(c, p.M, d);
      if (g) Gb[i] = G = g;
    }
    Gs[i] = G; waves[i] = w;
    pd.ftab = nullptr;
    if (b.arena && !Gb[i])
      return fail(c, MSVG_EINVAL, "allocation-free search: a profile's table does not fit the batched kernel "
                                  "(column-segmented models are searched by msvg_search)");
    if (Gb[i]) {
      const uint32_t *bt;
      if ((rc = ensure_bat_table(c, ids[i], dbid, G, &bt, st, b.arena)) != MSVG_OK) return rc;
      pd.ftab = reinterpret_cast<const uint4 *>(bt);
    } else if (!segmented[i]) {
      const uint32_t *ft;
      if ((rc = ensure_fast_table(c, ids[i], dbid, G, &ft, st)) != MSVG_OK) return rc;
      pd.ftab = reinterpret_cast<const uint4 *>(ft);
    } else if ((rc = ensure_seg_tables(c, ids[i], dbid, st)) != MSVG_OK) {
      return rc;
    }
    const uint32_t *stab;
    if ((rc = ensure_slow_table(c, ids[i], dbid, &stab, st, b.arena)) != MSVG_OK) return rc;
    const int Q64 = qfor(p.M, 128);
    jtab[i] = nullptr;
    if (Q64 <= kQMax && fast_lds(d.R, Q64) <= (size_t) kLdsBytes) {
      const uint32_t *jt;
      if (b.arena && Gb[i] == 64) jt = reinterpret_cast<const uint32_t *>(pd.ftab);   // the same table
      else if ((rc = ensure_fast_table(c, ids[i], dbid, 64, &jt, st, b.arena)) != MSVG_OK) return rc;
      jtab[i] = reinterpret_cast<const uint4 *>(jt);
    }
    pd.stab = stab;
    pd.M = p.M; pd.Q = qfor(p.M, 2 * G); pd.QB = (pd.Q + 3) / 4; pd.Qs = qfor(p.M, 128);
    pd.tbm = p.tbm; pd.tec = p.tec; pd.base = p.base; pd.bias = p.bias;
    pd.slow_all = 0;
    cells += (int64_t) p.M * d.total_res;
  }
  *cells_out = cells;
  if (tables_only) return MSVG_OK;          // the tables' builds are enqueued on st

  // Per-search buffers.
  HIPCHK(c, b.d_res.ensure((size_t) np * n * sizeof(uint16_t)));
  for (int i = 0; i < np; ++i) pds[i].res = b.d_res.as<uint16_t>() + (size_t) i * n;
  HIPCHK(c, b.d_prof.ensure(np * sizeof(ProfDev)));
  if (!capturing) HIPCHK(c, hipMemcpyAsync(b.d_prof.p, pds.data(), np * sizeof(ProfDev), hipMemcpyHostToDevice, st));
  if (b.arena) {
    b.slow_cap = noalloc_slow_cap(np, n);              // fixed: the scratch is sized for it
  } else {
    if (c->slow_cap < std::max<int64_t>(4096, (int64_t) np * n / 32))
      c->slow_cap = std::max<int64_t>(4096, (int64_t) np * n / 32);
    if (!capturing) b.slow_cap = c->slow_cap;
  }
  HIPCHK(c, b.d_slow.ensure((size_t) b.slow_cap * sizeof(SlowItem)));
  HIPCHK(c, b.d_counters.ensure((size_t) (2 * np + 2) * sizeof(int32_t)));
  if (!capturing) b.jcap = std::max<int64_t>(1024, n / 16);
  HIPCHK(c, b.d_jslow.ensure((size_t) np * b.jcap * sizeof(SlowItem)));
  if (!b.ext_out) {
    HIPCHK(c, b.d_usc.ensure((size_t) np * n * sizeof(float)));
    HIPCHK(c, b.d_status.ensure((size_t) np * n));
    HIPCHK(c, b.d_pass.ensure((size_t) np * n));
  }

  // Host-built finalize tables.
  std::vector<uint16_t> thr((size_t) np * d.Ls.size());
  std::vector<float> utab((size_t) np * 511);
  for (int i = 0; i < np; ++i) {
    const auto &ft = final_tables(c, ids[i], dbid, F1);
    std::copy(ft.first.begin(), ft.first.end(), thr.begin() + (size_t) i * d.Ls.size());
    std::copy(ft.second.begin(), ft.second.end(), utab.begin() + (size_t) i * 511);
  }
  HIPCHK(c, b.d_thr.ensure(thr.size() * sizeof(uint16_t) + 16));
  HIPCHK(c, b.d_usctab.ensure(utab.size() * sizeof(float)));
  if (!capturing) {
    HIPCHK(c, hipMemcpyAsync(b.d_thr.p, thr.data(), thr.size() * sizeof(uint16_t), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(b.d_usctab.p, utab.data(), utab.size() * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipStreamSynchronize(st));
  }

  DbDev dd;
  dd.res = d.d_res.as<uint8_t>(); dd.off = d.d_off.as<int64_t>(); dd.len = d.d_len.as<int32_t>();
  dd.tjb = d.d_tjb.as<uint8_t>(); dd.n = n; dd.R = d.R;
  int32_t *counters = b.d_counters.as<int32_t>();
  int32_t *slow_count = counters + np;

  if (timed) HIPCHK(c, hipEventRecord(c->ev[0], st));
  HIPCHK(c, msvk_fill_zero(counters, 2 * np + 2, st));
  int32_t *jcounts = counters + np + 1;
  int n_launch = 0;

  if (batch) {
    // One persistent launch per lane-group size G over the profiles it
    // takes: per-profile chunk queues, workgroups spread over the profiles
    // in proportion to M at the start, then moving to profiles with chunks
    // left (msv_batch.hip).
    struct BatLaunch { int G, grid, block; size_t lds, prof_at, first_at; int np; const int32_t *cs; int nchunks; };
    std::vector<BatLaunch> bl;
    std::vector<BatProf> hprof;
    std::vector<int32_t> hfirst;
    for (int G : {64, 32, 16, 8}) {
      std::vector<int> mem;
      for (int i = 0; i < np; ++i)
        if (Gb[i] == G) mem.push_back(i);
      if (mem.empty()) continue;
      const int32_t *cs;
      int nch;
      if ((rc = chunk_list(c, d, G, b.arena, &cs, &nch)) != MSVG_OK) return rc;
      size_t lds = 0;
      for (int i : mem) lds = std::max(lds, bat_bytes(G, d.R, pds[i].Q));
      lds += 16;                                        // the workgroup's broadcast word
      // Block size: the most resident waves per CU (LDS and VGPRs decide),
      // at most 16 waves per workgroup; ties go to the smaller block.
      int block = 0, occ = 0;
      for (int w : {4, 6, 8, 12, 16}) {
        int o = 0;
        HIPCHK(c, msvk_batch_occupancy(G, 64 * w, lds, &o));
        if (o * w > occ * (block / 64)) { block = 64 * w; occ = o; }
      }
      if (occ < 1) return fail(c, MSVG_EDEVICE, "batched kernel cannot be resident");
      BatLaunch L{G, c->ncu * occ, block, lds, hprof.size(), hfirst.size(), (int) mem.size(), cs, nch};
      double wsum = 0;
      for (int i : mem) wsum += c->profiles[ids[i]]->M;
      double acc = 0;
      size_t k = 0;
      for (int bidx = 0; bidx < L.grid; ++bidx) {
        const double at = (bidx + 0.5) / L.grid * wsum;
        while (k + 1 < mem.size() && acc + c->profiles[ids[mem[k]]]->M <= at) acc += c->profiles[ids[mem[k++]]]->M;
        hfirst.push_back((int32_t) k);
      }
      for (int i : mem) {
        BatProf bp;
        bp.tab = pds[i].ftab;
        bp.pd = pds[i];
        bp.pd.res = b.d_res.as<uint16_t>() + (size_t) i * n;
        bp.counter = counters + i;
        bp.jslow = jtab[i] ? b.d_jslow.as<SlowItem>() + (size_t) i * b.jcap : nullptr;
        bp.jslow_count = jcounts + i;
        bp.prof_idx = i;
        bp.tab_u4 = (int32_t) (bat_bytes(G, d.R, pds[i].Q) / 16);
        hprof.push_back(bp);
      }
      bl.push_back(L);
    }
    const size_t prof_bytes = (hprof.size() * sizeof(BatProf) + 255) & ~(size_t) 255;
    HIPCHK(c, b.d_bat.ensure(prof_bytes + hfirst.size() * sizeof(int32_t)));
    if (!capturing) {
      HIPCHK(c, hipMemcpyAsync(b.d_bat.p, hprof.data(), hprof.size() * sizeof(BatProf), hipMemcpyHostToDevice, st));
      HIPCHK(c, hipMemcpyAsync(b.d_bat.as<uint8_t>() + prof_bytes, hfirst.data(), hfirst.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, st));
      HIPCHK(c, hipStreamSynchronize(st));             // the host vectors end with this block
    }
    c->last_nbatch = 0;
    c->last_batch_timed = timed ? 1 : 0;
    for (const BatLaunch &L : bl) {
      const int k = c->last_nbatch++;
      c->last_batch_G[k] = L.G;
      c->last_batch_cells[k] = 0;
      for (int i = 0; i < np; ++i)
        if (Gb[i] == L.G) c->last_batch_cells[k] += (int64_t) c->profiles[ids[i]]->M * d.total_res;
      BatArgs a;
      a.profs = b.d_bat.as<BatProf>() + L.prof_at;
      a.np = L.np;
      a.db = dd;
      a.chunk_start = L.cs;
      a.nchunks = L.nchunks;
      a.first = reinterpret_cast<const int32_t *>(b.d_bat.as<uint8_t>() + prof_bytes) + L.first_at;
      a.slow = b.d_slow.as<SlowItem>();
      a.slow_count = slow_count;
      a.slow_cap = (int32_t) b.slow_cap;
      a.jslow_cap = (int32_t) b.jcap;
      a.slot_u4 = (int32_t) ((L.lds - 16) / 16);
      if (timed) HIPCHK(c, hipEventRecord(c->bev[2 * k], st));
      HIPCHK(c, msvk_launch_batch(L.G, &a, L.grid, L.block, L.lds, st));
      if (timed) HIPCHK(c, hipEventRecord(c->bev[2 * k + 1], st));
      ++n_launch;
    }
  }

  // Column-segmented profiles: one profile's segments run in order (each
  // reads the carries the one before wrote), different profiles need not.
  // With a side stream (batched searches) they go round-robin over up to
  // kJStreams streams forked from st, each stream with its own carry and
  // partial-maximum buffer, so one profile's launch tail overlaps the next
  // profile's work; otherwise in line on st.
  int nseg = 0;
  for (int i = 0; i < np; ++i) nseg += (!Gb[i] && segmented[i]) ? 1 : 0;
  const int nss = (side && nseg > 1) ? std::min(kJStreams, nseg) : 0;
  const size_t seg_bytes = ((size_t) d.packed * sizeof(uint16_t) + (size_t) d.n * sizeof(uint16_t) + 64 + 255) &
                           ~(size_t) 255;
  if (nseg) HIPCHK(c, b.d_seg.ensure((size_t) std::max(1, nss) * seg_bytes));
  if (nss) {
    HIPCHK(c, hipEventRecord(c->fork, st));
    for (int k = 0; k < nss; ++k) HIPCHK(c, hipStreamWaitEvent(c->jst[k], c->fork, 0));
  }
  int useg = 0;
  for (int i = 0; i < np; ++i) {
    const Profile &p = *c->profiles[ids[i]];
    if (Gb[i]) continue;                                // scored by its batched launch
    if (segmented[i]) {
      // Long model: column segments of kSegQ registers, carrying the last
      // column of each row from one segment to the next.
      Tables &t = tables_for(c, ids[i], dbid);
      const hipStream_t ss = nss ? c->jst[useg % nss] : st;
      uint8_t *segbuf = b.d_seg.as<uint8_t>() + (size_t) (nss ? useg % nss : 0) * seg_bytes;
      ++useg;
      const int span = 128 * seg_q(d.R);
      int si = 0;
      for (int k0 = 0; k0 < p.M; k0 += span, ++si) {
        const int Mseg = std::min(span, p.M - k0);
        const int Q = qfor(Mseg, 128);
        SegArgs a;
        a.prof = pds[i];
        a.prof.Q = Q; a.prof.QB = (Q + 3) / 4;
        a.db = dd;
        a.tab = t.seg[si]->as<uint4>();
        a.carry = reinterpret_cast<uint16_t *>(segbuf);
        a.partial = reinterpret_cast<uint16_t *>(segbuf + (size_t) d.packed * sizeof(uint16_t));
        a.first = (k0 == 0); a.last = (k0 + span >= p.M);
        a.prof_idx = i; a.slow = b.d_slow.as<SlowItem>(); a.slow_count = slow_count; a.slow_cap = (int32_t) b.slow_cap;
        a.jslow = jtab[i] ? b.d_jslow.as<SlowItem>() + (size_t) i * b.jcap : nullptr;
        a.jslow_count = jcounts + i;
        a.jslow_cap = (int32_t) b.jcap;
        a.carry_base = d.d_off.as<int64_t>();
        const size_t lds = fast_lds(d.R, Q);
        int grid, block;
        persistent_shape(c, lds, 0, grid, block);
        a.ntasks = (int32_t) n;
        HIPCHK(c, msvk_launch_seg(Q, &a, grid, block, lds, ss));
        ++n_launch;
      }
      if (jtab[i]) {
        JArgs ja;
        ja.prof = pds[i];
        ja.db = dd;
        ja.tab = jtab[i];
        ja.items = b.d_jslow.as<SlowItem>() + (size_t) i * b.jcap;
        ja.count = jcounts + i;
        ja.cap = (int32_t) b.jcap;
        const int Q64 = qfor(p.M, 128);
        const size_t jl = fast_lds(d.R, Q64);
        const int per_cu = (int) std::max<size_t>(1, std::min<size_t>(4, kLdsBytes / jl));
        if (nss) HIPCHK(c, msvk_launch_jrow(Q64, &ja, c->ncu * per_cu, jl, ss));   // after its last segment
        else if ((rc = launch_j(ja, Q64, jl, c->ncu * per_cu)) != MSVG_OK) return rc;
      }
      continue;
    }
    FastArgs a;
    a.prof = pds[i];
    a.db = dd;
    a.prof_idx = i;
    a.slow = b.d_slow.as<SlowItem>();
    a.slow_count = slow_count;
    a.slow_cap = (int32_t) b.slow_cap;
    a.counter = counters + i;
    a.jslow = jtab[i] ? b.d_jslow.as<SlowItem>() + (size_t) i * b.jcap : nullptr;
    a.jslow_count = jcounts + i;
    a.jslow_cap = (int32_t) b.jcap;
    const int G = Gs[i];
    const size_t lds = fast_lds(d.R, pds[i].Q);
    int grid, block;
    if (mode == MSVG_MODE_BATCHED) {
      if ((rc = build_chunks(c, d, G)) != MSVG_OK) return rc;
      auto &ch = d.chunks[G];
      a.mode = sched == 2 ? 2 : 0;
      a.chunk_start = ch.first->as<int32_t>();
      a.nchunks = ch.second;
      a.order = nullptr; a.ntasks = 0;
      persistent_shape(c, lds, waves[i], grid, block);
      if (sched == 2) { grid = std::max(1, ch.second); block = 256; }
    } else {
      a.mode = 1;
      a.chunk_start = nullptr; a.nchunks = 0;
      a.order = d.d_order.as<int32_t>();
      a.ntasks = (int32_t) n;
      block = 256;
      grid = c->ncu * (int) std::max<size_t>(1, std::min<size_t>(8, kLdsBytes / lds));
    }
    HIPCHK(c, msvk_launch_fast(G, pds[i].Q, &a, grid, block, lds, st));
    ++n_launch;
    if (jtab[i]) {
      JArgs ja;
      ja.prof = pds[i];
      ja.db = dd;
      ja.tab = jtab[i];
      ja.items = a.jslow;
      ja.count = a.jslow_count;
      ja.cap = a.jslow_cap;
      const int Q64 = qfor(p.M, 128);
      const size_t jl = fast_lds(d.R, Q64);
      const int per_cu = (int) std::max<size_t>(1, std::min<size_t>(4, kLdsBytes / jl));
      if ((rc = launch_j(ja, Q64, jl, c->ncu * per_cu)) != MSVG_OK) return rc;
    }
  }
  if (nss)
    for (int k = 0; k < nss; ++k) {
      HIPCHK(c, hipEventRecord(c->jdone[k], c->jst[k]));
      HIPCHK(c, hipStreamWaitEvent(st, c->jdone[k], 0));
    }
  if (side_used) {
    HIPCHK(c, hipEventRecord(c->join, side));
    HIPCHK(c, hipStreamWaitEvent(st, c->join, 0));
  }
  if (batch) {
    // The J-aware pairs of the batched profiles, once the batched launches
    // have filled their queues: one msv_jrow_many launch, every profile's
    // pairs side by side.  (One launch per profile lasts as long as its
    // longest pair's row chain, so even on four streams the launches after
    // the first on each stream add their whole chains.)
    std::vector<JProf> hjp;
    size_t jlds = 0;
    for (int i = 0; i < np; ++i) {
      if (!Gb[i] || !jtab[i]) continue;
      JProf jp{};
      jp.prof = pds[i];
      jp.prof.res = b.d_res.as<uint16_t>() + (size_t) i * n;
      jp.tab = jtab[i];
      jp.items = b.d_jslow.as<SlowItem>() + (size_t) i * b.jcap;
      jp.count = jcounts + i;
      jp.cap = (int32_t) b.jcap;
      jp.q = qfor(c->profiles[ids[i]]->M, 128);
      jlds = std::max(jlds, fast_lds(d.R, jp.q));
      hjp.push_back(jp);
    }
    if (timed) HIPCHK(c, hipEventRecord(c->bev[8], st));
    if (!hjp.empty()) {
      HIPCHK(c, b.d_jprof.ensure(hjp.size() * sizeof(JProf)));
      if (!capturing) {
        HIPCHK(c, hipMemcpyAsync(b.d_jprof.p, hjp.data(), hjp.size() * sizeof(JProf), hipMemcpyHostToDevice, st));
        HIPCHK(c, hipStreamSynchronize(st));           // hjp ends with this block
      }
      int per_cu = 0;
      HIPCHK(c, msvk_jrow_many_occupancy(jlds, &per_cu));
      JManyArgs ja;
      ja.db = dd;
      ja.profs = b.d_jprof.as<JProf>();
      ja.np = (int32_t) hjp.size();
      // about twice the workgroups the device holds, each profile the same
      // number (its workgroups beyond its pairs leave at once)
      ja.wpp = std::max(1, (2 * c->ncu * std::max(1, per_cu) + ja.np - 1) / ja.np);
      HIPCHK(c, msvk_launch_jrow_many(&ja, jlds, st));
    }
    if (timed) HIPCHK(c, hipEventRecord(c->bev[9], st));
  }
  if (timed) HIPCHK(c, hipEventRecord(c->ev[1], st));
  c->last_launches = n_launch;

  SlowArgs sa;
  sa.profs = b.d_prof.as<ProfDev>();
  sa.db = dd;
  sa.items = b.d_slow.as<SlowItem>();
  sa.count = slow_count;
  sa.cap = (int32_t) b.slow_cap;
  HIPCHK(c, msvk_launch_slow(&sa, c->ncu * 4, st));
  if (timed) HIPCHK(c, hipEventRecord(c->ev[2], st));

  FinArgs fa;
  fa.profs = b.d_prof.as<ProfDev>();
  fa.db = dd;
  fa.sorted_pos = d.d_sorted_pos.as<int64_t>();
  fa.lidx = d.d_lidx.as<int32_t>();
  fa.usc_tab = b.d_usctab.as<float>();
  fa.thr = b.d_thr.as<uint16_t>();
  fa.nL = (int32_t) d.Ls.size();
  fa.usc = b.d_usc.as<float>();
  fa.status = b.d_status.as<uint8_t>();
  fa.pass = b.d_pass.as<uint8_t>();
  HIPCHK(c, msvk_launch_finalize(&fa, np, n, st));
  if (timed) HIPCHK(c, hipEventRecord(c->ev[3], st));
  return MSVG_OK;
}

// Scratch bytes an allocation-free search of these profiles over database
// dbid takes (the arena's 256-byte alignment included); 0 if a profile is
// not taken by the batched kernel.
size_t noalloc_bytes(msvg_ctx *c, const int32_t *ids, int np, int dbid) {
  const Db &d = *c->dbs[dbid];
  const int64_t n = d.n;
  size_t tot = 0;
  auto add = [&](size_t b) { tot += ((b + 255) & ~(size_t) 255) + 256; };
  std::vector<int> Gs;
  for (int i = 0; i < np; ++i) {
    const int M = c->profiles[ids[i]]->M;
    const int G = batch_lanes(c, M, d);
    if (!G) return 0;
    add(bat_bytes(G, d.R, qfor(M, 2 * G)));
    const int Q64 = qfor(M, 128);
    if (G != 64 && Q64 <= kQMax && fast_lds(d.R, Q64) <= (size_t) kLdsBytes) add(fast_lds(d.R, Q64));
    add((size_t) (d.R + 1) * Q64 * 64 * 4);                          // byte-stage table
    if (std::find(Gs.begin(), Gs.end(), G) == Gs.end()) Gs.push_back(G);
  }
  for (int G : Gs) add(guided_chunks(c, d, G).size() * sizeof(int32_t));
  add((size_t) np * n * sizeof(uint16_t));                            // result words
  add((size_t) np * sizeof(ProfDev));
  add((size_t) noalloc_slow_cap(np, n) * sizeof(SlowItem));
  add((size_t) (2 * np + 2) * sizeof(int32_t));
  add((size_t) np * sizeof(JProf));
  add((size_t) np * std::max<int64_t>(1024, n / 16) * sizeof(SlowItem));
  add((size_t) np * d.Ls.size() * sizeof(uint16_t) + 16);
  add((size_t) np * 511 * sizeof(float));
  add(((np * sizeof(BatProf) + 255) & ~(size_t) 255) + Gs.size() * (size_t) c->ncu * 16 * sizeof(int32_t));
  return tot;
}

int after_search(msvg_ctx *c, Bufs &b, int np, int dbid, int64_t cells, bool timed) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> cn(2 * np + 2);
  HIPCHK(c, hipMemcpy(cn.data(), b.d_counters.p, cn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  const int32_t cnt = cn[np];
  int64_t jtot = 0;
  for (int i = 0; i < np; ++i) jtot += std::min<int64_t>(cn[np + 1 + i], b.jcap);
  c->last = &b;
  c->stats.cells = cells;
  c->stats.n_slow = cnt;
  c->stats.n_jrow = jtot;
  c->stats.dropin_path = 0;
  if (timed) {
    float a = 0, b = 0, f = 0, t = 0;
    HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIPCHK(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    HIPCHK(c, hipEventElapsedTime(&f, c->ev[2], c->ev[3]));
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
    c->stats.ms_fast = a; c->stats.ms_slow = b; c->stats.ms_finalize = f; c->stats.ms_total = t;
  }
  c->stats.n_batch = c->last_nbatch;
  c->stats.ms_jrow = 0;
  for (int k = 0; k < 4; ++k) {
    c->stats.batch_lanes[k] = k < c->last_nbatch ? c->last_batch_G[k] : 0;
    c->stats.batch_cells[k] = k < c->last_nbatch ? c->last_batch_cells[k] : 0;
    c->stats.batch_ms[k] = 0;
  }
  if (timed && c->last_batch_timed && c->last_nbatch > 0) {
    for (int k = 0; k < c->last_nbatch; ++k) {
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, c->bev[2 * k], c->bev[2 * k + 1]));
      c->stats.batch_ms[k] = ms;
    }
    float mj = 0;
    HIPCHK(c, hipEventElapsedTime(&mj, c->bev[8], c->bev[9]));
    c->stats.ms_jrow = mj;
  }
  c->last_np = np;
  c->last_db = dbid;
  if (cnt > b.slow_cap) return MSVG_ENORESULT;     // caller re-runs with a larger queue
  return MSVG_OK;
}

int search_device(msvg_ctx *c, const int32_t *ids, int np, int dbid, double F1, int mode_flags) {
  int rc = check_ids(c, ids, np, dbid);
  if (rc != MSVG_OK) return rc;
  const bool eager = (mode_flags & MSVG_FLAG_EAGER) != 0;
  const int mode = mode_flags & 0xf;
  for (int attempt = 0; attempt < 3; ++attempt) {
    int64_t cells = 0;
    if (mode == MSVG_MODE_BATCHED && !eager) {
      std::string key = std::to_string(dbid) + ":" + std::to_string(F1) + ":" + std::to_string(c->slow_cap);
      for (int i = 0; i < np; ++i) key += "," + std::to_string(ids[i]);
      auto &plan = c->plans[key];
      if (!plan) {
        // Warm every cache (tables, chunks, buffers) outside the capture, then
        // capture the launch sequence once.
        auto pl = std::make_unique<Plan>();
        if ((rc = enqueue_search(c, pl->bufs, ids, np, dbid, F1, mode, false, false, &cells, nullptr, 0,
                                 c->side)) != MSVG_OK) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        rc = enqueue_search(c, pl->bufs, ids, np, dbid, F1, mode, false, true, &cells, nullptr, 0, c->side);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != MSVG_OK) { if (g) (void) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
        pl->graph = g;
        HIPCHK(c, hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0));
        plan = std::move(pl);
      }
      for (int i = 0; i < np; ++i) cells += (int64_t) c->profiles[ids[i]]->M * c->dbs[dbid]->total_res;
      HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
      HIPCHK(c, hipGraphLaunch(plan->exec, c->stream));
      HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      float t = 0;
      HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
      rc = after_search(c, plan->bufs, np, dbid, cells, false);
      c->stats.ms_total = t; c->stats.ms_fast = c->stats.ms_slow = c->stats.ms_finalize = 0;
      c->stats.n_fast_launches = c->last_launches;
    } else {
      if ((rc = enqueue_search(c, c->eager, ids, np, dbid, F1, mode, true, false, &cells, nullptr, 0,
                               mode == MSVG_MODE_BATCHED ? c->side : nullptr)) != MSVG_OK) return rc;
      rc = after_search(c, c->eager, np, dbid, cells, true);
      c->stats.n_fast_launches = c->last_launches;
    }
    if (rc != MSVG_ENORESULT) return rc;
    c->slow_cap = (int64_t) c->stats.n_slow + 4096;     // queue overflowed: grow and redo
    c->last = nullptr;
    c->plans.clear();
  }
  return fail(c, MSVG_EMEM, "slow-path queue could not be sized");
}

}  // namespace

// ViterbiFilter tables (host-built, see VitArgs in msv_device.h).
namespace {

// Registers per lane of the one-wave vit_band shape beyond 16 (MM, IM and DM
// in LDS): as many as keep its tables within 160 KiB of LDS for a database
// holding all Kp codes (Kp + 1 emission rows and the three transition rows,
// 64 lanes x 4 bytes each), at most 20 -- 19 for amino, 20 for DNA.
int vit_band64_qmax(int Kp) { return std::min(20, (160 * 1024) / (256 * (Kp + 4))); }

// ViterbiFilter kernel shape for a model of length M over an alphabet of Kp
// codes (measured on gfx950, survivors of up to 2,000 residues): vit_band
// with G = 16 lanes per target (four targets per wave) up to M = 512, G = 32
// up to 1024, G = 64 up to 128 x vit_band64_qmax(Kp) (Q = ceil(M / 2G));
// beyond, vit_kernel with W = 2, 4 or 8 waves per target (Q <= 16).
void vit_shape(int M, int Kp, int &Q, int &W, int &G) {
  W = 1;
  G = (M <= 512) ? 16 : (M <= 1024 ? 32 : 64);
  if (M > 128 * vit_band64_qmax(Kp)) while (128 * W * 16 < M) W *= 2;
  Q = (M + 2 * G * W - 1) / (2 * G * W);
}

// Lanes per target of a shape.
int vit_lanes(const VitProfile &v) { return v.G < 64 ? v.G : 64 * v.W; }

// Pack two int16 model-order values for lane c, register j (slots 2c, 2c+1)
// of a target striped over C lanes.
template <class F>
void vit_fill(std::vector<uint32_t> &out, size_t base, int Q, int C, int M, F value) {
  for (int j = 0; j < Q; ++j)
    for (int c = 0; c < C; ++c) {
      const int k0 = j + (2 * c) * Q + 1, k1 = j + (2 * c + 1) * Q + 1;
      const int v0 = k0 <= M ? value(k0) : -32768, v1 = k1 <= M ? value(k1) : -32768;
      out[base + (size_t) j * C + c] = (uint32_t) (uint16_t) v0 | ((uint32_t) (uint16_t) v1 << 16);
    }
}

int vit_emission_table(msvg_ctx *c, int pid, int dbid, const uint32_t **out) {
  Profile &p = *c->profiles[pid];
  VitProfile &v = *p.vit;
  Db &d = *c->dbs[dbid];
  Tables &t = tables_for(c, pid, dbid);
  if (!t.vemit) {
    const size_t per = (size_t) v.Q * vit_lanes(v);
    std::vector<uint32_t> h((size_t) (d.R + 1) * per);
    for (int r = 0; r <= d.R; ++r) {
      const int x = (r < d.R) ? d.rowcode[r] : -1;
      vit_fill(h, (size_t) r * per, v.Q, vit_lanes(v), p.M,
               [&](int k) { return x < 0 ? -32768 : (int) v.rwv[(size_t) x * p.M + k - 1]; });
    }
    auto mem = std::make_unique<DevMem>();
    HIPCHK(c, mem->ensure(h.size() * 4));
    HIPCHK(c, hipMemcpy(mem->p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    t.vemit = std::move(mem);
  }
  *out = t.vemit->as<uint32_t>();
  return MSVG_OK;
}

// The column segments' emission tables (vit_seg; 64-lane layout) over the
// database's rows.
int vit_seg_emission_tables(msvg_ctx *c, int pid, int dbid) {
  Profile &p = *c->profiles[pid];
  VitProfile &v = *p.vit;
  Db &d = *c->dbs[dbid];
  Tables &t = tables_for(c, pid, dbid);
  if (!t.vseg_emit.empty()) return MSVG_OK;
  for (int k0 = 0; k0 < p.M; k0 += v.seg_span) {
    const int Ms = std::min(v.seg_span, p.M - k0), Qs = (Ms + 127) / 128;
    const size_t per = (size_t) Qs * 64;
    std::vector<uint32_t> h((size_t) (d.R + 1) * per);
    for (int r = 0; r <= d.R; ++r) {
      const int x = (r < d.R) ? d.rowcode[r] : -1;
      vit_fill(h, (size_t) r * per, Qs, 64, Ms,
               [&](int k) { return x < 0 ? -32768 : (int) v.rwv[(size_t) x * p.M + k0 + k - 1]; });
    }
    auto mem = std::make_unique<DevMem>();
    HIPCHK(c, mem->ensure(h.size() * 4));
    HIPCHK(c, hipMemcpy(mem->p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    t.vseg_emit.push_back(std::move(mem));
  }
  return MSVG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int msvg_init(int device, msvg_ctx **ret) {
  if (!ret) return MSVG_EINVAL;
  *ret = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MSVG_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return MSVG_EDEVICE;
  auto c = std::make_unique<msvg_ctx>();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MSVG_EDEVICE;
  c->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return MSVG_EDEVICE;
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) return MSVG_EDEVICE;
  if (hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess) return MSVG_EDEVICE;
  if (hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess) return MSVG_EDEVICE;
  for (int k = 0; k < kJStreams; ++k) {
    if (hipStreamCreateWithFlags(&c->jst[k], hipStreamNonBlocking) != hipSuccess) return MSVG_EDEVICE;
    if (hipEventCreateWithFlags(&c->jdone[k], hipEventDisableTiming) != hipSuccess) return MSVG_EDEVICE;
  }
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return MSVG_EDEVICE;
  for (auto &e : c->bev)
    if (hipEventCreate(&e) != hipSuccess) return MSVG_EDEVICE;
  if (msvk_prepare() != hipSuccess) return MSVG_EDEVICE;
  if (msvk_prepare_batch() != hipSuccess) return MSVG_EDEVICE;
  if (msvk_prepare_scan_group() != hipSuccess) return MSVG_EDEVICE;
  if (hipGetSymbolAddress((void **) &c->scratch_dev, HIP_SYMBOL(scratch)) != hipSuccess || !c->scratch_dev)
    return MSVG_EDEVICE;
  *ret = c.release();
  return MSVG_OK;
}

void msvg_free(msvg_ctx *c) {
  if (!c) return;
  (void) hipSetDevice(c->device);
  (void) hipStreamSynchronize(c->stream);
  c->plans.clear();
  c->tables.clear();
  c->dbs.clear();
  c->profiles.clear();
  for (auto &e : c->ev) if (e) (void) hipEventDestroy(e);
  for (auto &e : c->bev) if (e) (void) hipEventDestroy(e);
  if (c->fork) (void) hipEventDestroy(c->fork);
  if (c->join) (void) hipEventDestroy(c->join);
  if (c->side) { (void) hipStreamSynchronize(c->side); (void) hipStreamDestroy(c->side); }
  for (int k = 0; k < kJStreams; ++k) {
    if (c->jst[k]) { (void) hipStreamSynchronize(c->jst[k]); (void) hipStreamDestroy(c->jst[k]); }
    if (c->jdone[k]) (void) hipEventDestroy(c->jdone[k]);
  }
  for (int k = 0; k < kVitStreams; ++k) {
    if (c->vstream[k]) { (void) hipStreamSynchronize(c->vstream[k]); (void) hipStreamDestroy(c->vstream[k]); }
    if (c->vdone[k]) (void) hipEventDestroy(c->vdone[k]);
  }
  if (c->stream) (void) hipStreamDestroy(c->stream);
  if (c->h_one_res) (void) hipHostFree(c->h_one_res);
  if (c->h_one_out) (void) hipHostFree(c->h_one_out);
  if (c->h_scanout) (void) hipHostFree(c->h_scanout);
  c->scanset.reset();
  delete c;
}

const char *msvg_last_error(const msvg_ctx *c) { return c ? c->err.c_str() : "null context"; }

void msvg_test_alloc_budget(long k) { g_alloc_budget.store(k); }

int msvg_profile_add(msvg_ctx *c, const msvg_oprofile *om, int32_t *ret_id) {
  if (!c || !om || !ret_id || !om->rbv) return fail(c, MSVG_EINVAL, "null argument");
  if (om->M < 1 || om->M > 16384) return fail(c, MSVG_EINVAL, "M must be in 1..16384");
  if (om->Kp < 1 || om->Kp > 255) return fail(c, MSVG_EINVAL, "Kp out of range");
  if ((int) om->bias_b + (int) om->base_b >= 255)
    return fail(c, MSVG_EINVAL, "bias_b + base_b >= 255 is outside the supported range");
  (void) hipSetDevice(c->device);
  auto p = std::make_unique<Profile>();
  p->M = om->M; p->Kp = om->Kp;
  p->rbv.assign(om->rbv, om->rbv + (size_t) om->M * om->Kp);
  p->tbm = om->tbm_b; p->tec = om->tec_b; p->base = om->base_b; p->bias = om->bias_b;
  p->scale = om->scale_b; p->mu = om->mmu; p->lambda = om->mlambda;
  HIPCHK(c, p->d_rbv.ensure(p->rbv.size()));
  HIPCHK(c, hipMemcpy(p->d_rbv.p, p->rbv.data(), p->rbv.size(), hipMemcpyHostToDevice));
  const int id = c->next_id++;
  c->profiles[id] = std::move(p);
  *ret_id = id;
  return MSVG_OK;
}

int msvg_profile_remove(msvg_ctx *c, int32_t id) {
  if (!c || !c->profiles.count(id)) return fail(c, MSVG_EINVAL, "unknown profile id");
  (void) hipStreamSynchronize(c->stream);
  for (auto it = c->tables.begin(); it != c->tables.end();)
    it = (it->first.first == id) ? c->tables.erase(it) : std::next(it);
  for (auto it = c->finals.begin(); it != c->finals.end();)
    it = (std::get<0>(it->first) == id) ? c->finals.erase(it) : std::next(it);
  c->last = nullptr;
  c->plans.clear();
  c->scanset.reset();
  c->profiles.erase(id);
  return MSVG_OK;
}

int msvg_db_add(msvg_ctx *c, const uint8_t *res, const int64_t *off, const int32_t *len, int64_t n, int32_t *ret_id) {
  if (!c || !ret_id || n < 0 || (n > 0 && (!res || !off || !len))) return fail(c, MSVG_EINVAL, "null argument");
  if (n >= (int64_t) 1 << 31) return fail(c, MSVG_EINVAL, "too many targets for one database");
  (void) hipSetDevice(c->device);
  auto d = std::make_unique<Db>();
  d->n = n;
  // The passes over the residues run on host threads over disjoint ranges
  // of targets (a large database is hundreds of MB).
  const int nth = (int) std::max<int64_t>(1, std::min<int64_t>({(int64_t) 16, n / 16384 + 1,
                                                                 (int64_t) std::max(1u, std::thread::hardware_concurrency())}));
  auto par = [nth](int64_t count, const std::function<void(int, int64_t, int64_t)> &fn) {
    std::vector<std::thread> th;
    int t = 1;
    try {
      for (; t < nth; ++t) th.emplace_back(fn, t, count * t / nth, count * (t + 1) / nth);
    } catch (...) {                    // a thread that cannot start: its ranges run here
    }
    for (int u = t; u < nth; ++u) fn(u, count * u / nth, count * (u + 1) / nth);
    fn(0, 0, count / nth);
    for (auto &x : th) x.join();
  };
  // Codes present, total residues, longest target.
  std::vector<std::array<uint8_t, 256>> seen_t((size_t) nth);
  std::vector<int64_t> total_t((size_t) nth, 0);
  std::vector<int32_t> maxl_t((size_t) nth, 0);
  std::vector<char> neg_t((size_t) nth, 0);
  par(n, [&](int t, int64_t lo, int64_t hi) {
    auto &sn = seen_t[(size_t) t];
    sn.fill(0);
    int64_t tot = 0;
    int32_t mx = 0;
    for (int64_t i = lo; i < hi; ++i) {
      if (len[i] < 0) { neg_t[(size_t) t] = 1; return; }
      const uint8_t *q = res + off[i];
      for (int32_t k = 0; k < len[i]; ++k) sn[q[k]] = 1;
      tot += len[i];
      mx = std::max(mx, len[i]);
    }
    total_t[(size_t) t] = tot;
    maxl_t[(size_t) t] = mx;
  });
  bool seen[256] = {false};
  int64_t total = 0;
  int32_t maxlen = 0;
  for (int t = 0; t < nth; ++t) {
    if (neg_t[(size_t) t]) return fail(c, MSVG_EINVAL, "negative target length");
    for (int x = 0; x < 256; ++x) seen[x] = seen[x] || seen_t[(size_t) t][(size_t) x];
    total += total_t[(size_t) t];
    maxlen = std::max(maxlen, maxl_t[(size_t) t]);
  }
  uint8_t rowof[256];
  for (int x = 0; x < 256; ++x)
    if (seen[x]) { rowof[x] = (uint8_t) d->rowcode.size(); d->rowcode.push_back((uint8_t) x); d->maxcode = x; }
  if (d->rowcode.empty()) d->rowcode.push_back(0);
  d->R = (int) d->rowcode.size();
  if (d->R > 254) return fail(c, MSVG_EINVAL, "too many distinct residue codes");
  d->total_res = total;
  // Length-sorted order (descending; ties by caller index for determinism),
  // with the distinct lengths (ascending) and each target's index among them.
  std::vector<int64_t> order(n);
  d->sorted_of.resize(n);
  d->len_sorted.resize(n);
  d->tjb_sorted.resize(n);
  d->lidx_sorted.resize(n);
  if (maxlen <= (1 << 20)) {
    // counting sort on the length: stable, so ties keep the caller's order
    std::vector<int64_t> start((size_t) maxlen + 2, 0);
    for (int64_t i = 0; i < n; ++i) start[(size_t) len[i]]++;
    std::vector<int32_t> idx_of((size_t) maxlen + 1, -1);
    std::vector<uint8_t> tjb_of((size_t) maxlen + 1, 0);
    for (int32_t L = 0; L <= maxlen; ++L)
      if (start[(size_t) L]) {
        idx_of[(size_t) L] = (int32_t) d->Ls.size();
        d->Ls.push_back(L);
        tjb_of[(size_t) L] = (uint8_t) msvh_tjb(L);
      }
    int64_t acc = 0;
    for (int32_t L = maxlen; L >= 0; --L) {
      const int64_t k = start[(size_t) L];
      start[(size_t) L] = acc;
      acc += k;
    }
    for (int64_t i = 0; i < n; ++i) order[(size_t) start[(size_t) len[i]]++] = i;
    for (int64_t sp = 0; sp < n; ++sp) {
      const int32_t L = len[order[(size_t) sp]];
      d->len_sorted[sp] = L;
      d->tjb_sorted[sp] = tjb_of[(size_t) L];
      d->lidx_sorted[sp] = idx_of[(size_t) L];
    }
  } else {
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return len[a] > len[b]; });
    for (int64_t sp = 0; sp < n; ++sp) {
      d->len_sorted[sp] = len[order[sp]];
      d->tjb_sorted[sp] = (uint8_t) msvh_tjb(d->len_sorted[sp]);
    }
    std::vector<int32_t> ls(d->len_sorted);
    std::sort(ls.begin(), ls.end());
    ls.erase(std::unique(ls.begin(), ls.end()), ls.end());
    d->Ls = ls;
    for (int64_t sp = 0; sp < n; ++sp)
      d->lidx_sorted[sp] = (int32_t) (std::lower_bound(ls.begin(), ls.end(), d->len_sorted[sp]) - ls.begin());
  }
  std::vector<int64_t> off_sorted(n);
  int64_t bytes = 0;
  for (int64_t sp = 0; sp < n; ++sp) {
    d->sorted_of[order[sp]] = sp;
    off_sorted[sp] = bytes;
    bytes += ((int64_t) d->len_sorted[sp] + 15) / 16 * 16;
  }
  d->packed = bytes + 64;
  // Remapped residues in sorted order, each target padded to 16 with the
  // null row R, 64 more bytes of R at the end.
  std::unique_ptr<uint8_t[]> packed(new uint8_t[(size_t) bytes + 64]);
  std::memset(packed.get() + bytes, d->R, 64);
  par(n, [&](int, int64_t lo, int64_t hi) {
    for (int64_t sp = lo; sp < hi; ++sp) {
      const int64_t i = order[sp];
      const int32_t L = len[i];
      const uint8_t *src = res + off[i];
      uint8_t *dst = packed.get() + off_sorted[sp];
      for (int32_t k = 0; k < L; ++k) dst[k] = rowof[src[k]];
      std::memset(dst + L, d->R, (size_t) ((((int64_t) L + 15) & ~(int64_t) 15) - L));
    }
  });
  std::vector<int32_t> order32(n);
  for (int64_t i = 0; i < n; ++i) order32[i] = (int32_t) d->sorted_of[i];

  HIPCHK(c, d->d_res.ensure((size_t) bytes + 64));
  HIPCHK(c, hipMemcpy(d->d_res.p, packed.get(), (size_t) bytes + 64, hipMemcpyHostToDevice));
  HIPCHK(c, d->d_off.ensure(std::max<int64_t>(n, 1) * 8));
  HIPCHK(c, hipMemcpy(d->d_off.p, off_sorted.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, d->d_len.ensure(std::max<int64_t>(n, 1) * 4));
  HIPCHK(c, hipMemcpy(d->d_len.p, d->len_sorted.data(), n * 4, hipMemcpyHostToDevice));
  HIPCHK(c, d->d_tjb.ensure(std::max<int64_t>(n, 1)));
  HIPCHK(c, hipMemcpy(d->d_tjb.p, d->tjb_sorted.data(), n, hipMemcpyHostToDevice));
  HIPCHK(c, d->d_lidx.ensure(std::max<int64_t>(n, 1) * 4));
  HIPCHK(c, hipMemcpy(d->d_lidx.p, d->lidx_sorted.data(), n * 4, hipMemcpyHostToDevice));
  HIPCHK(c, d->d_sorted_pos.ensure(std::max<int64_t>(n, 1) * 8));
  HIPCHK(c, hipMemcpy(d->d_sorted_pos.p, d->sorted_of.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHK(c, d->d_order.ensure(std::max<int64_t>(n, 1) * 4));
  HIPCHK(c, hipMemcpy(d->d_order.p, order32.data(), n * 4, hipMemcpyHostToDevice));
  HIPCHK(c, d->d_rowcode.ensure(d->rowcode.size()));
  HIPCHK(c, hipMemcpy(d->d_rowcode.p, d->rowcode.data(), d->rowcode.size(), hipMemcpyHostToDevice));
  const int id = c->next_id++;
  c->dbs[id] = std::move(d);
  *ret_id = id;
  return MSVG_OK;
}

int msvg_db_remove(msvg_ctx *c, int32_t id) {
  if (!c || !c->dbs.count(id)) return fail(c, MSVG_EINVAL, "unknown database id");
  (void) hipStreamSynchronize(c->stream);
  for (auto it = c->tables.begin(); it != c->tables.end();)
    it = (it->first.second == id) ? c->tables.erase(it) : std::next(it);
  for (auto it = c->finals.begin(); it != c->finals.end();)
    it = (std::get<1>(it->first) == id) ? c->finals.erase(it) : std::next(it);
  c->last = nullptr;
  c->plans.clear();
  c->dbs.erase(id);
  if (c->last_db == id) c->last_db = -1;
  return MSVG_OK;
}

int64_t msvg_db_size(msvg_ctx *c, int32_t id) {
  if (!c || !c->dbs.count(id)) return -1;
  return c->dbs[id]->n;
}

int msvg_search_device(msvg_ctx *c, const int32_t *ids, int32_t np, int32_t dbid, double F1, int mode) {
  if (!c) return MSVG_EINVAL;
  if ((mode & 0xf) != MSVG_MODE_PLAIN && (mode & 0xf) != MSVG_MODE_BATCHED) return fail(c, MSVG_EINVAL, "unknown mode");
  (void) hipSetDevice(c->device);
  return search_device(c, ids, np, dbid, F1, mode);
}

int msvg_search_noalloc(msvg_ctx *c, const int32_t *ids, int32_t np, int32_t dbid, double F1, float *d_usc,
                        uint8_t *d_status, uint8_t *d_pass) {
  int rc = check_ids(c, ids, np, dbid);
  if (rc != MSVG_OK) return rc;
  (void) hipSetDevice(c->device);
  const size_t need = noalloc_bytes(c, ids, np, dbid);
  if (need == 0)
    return fail(c, MSVG_EINVAL, "allocation-free search: a profile's table does not fit the batched kernel "
                                "(column-segmented models are searched by msvg_search)");
  if (need > kScratchBytes)
    return fail(c, MSVG_EINVAL, "allocation-free search: " + std::to_string(np) + " profiles x " +
                                    std::to_string(c->dbs[dbid]->n) + " targets need " + std::to_string(need) +
                                    " bytes of scratch; the limit is " + std::to_string(kScratchBytes) +
                                    " (search fewer profiles or targets per call)");
  Arena ar;
  ar.base = c->scratch_dev;
  ar.cap = kScratchBytes;
  Bufs b(&ar);
  b.ext_out = true;                            // results go straight to the caller's arrays
  for (DevMem *m : {&b.d_usc, &b.d_status, &b.d_pass}) { m->arena = nullptr; m->borrowed = true; }
  b.d_usc.p = d_usc;
  b.d_status.p = d_status;
  b.d_pass.p = d_pass;
  int64_t cells = 0;
  rc = enqueue_search(c, b, ids, np, dbid, F1, MSVG_MODE_BATCHED, false, false, &cells, c->stream, 0, c->side);
  if (rc != MSVG_OK) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int32_t n_slow = 0;
  HIPCHK(c, hipMemcpy(&n_slow, b.d_counters.as<int32_t>() + np, sizeof(int32_t), hipMemcpyDeviceToHost));
  c->stats = msvg_stats{};
  c->stats.cells = cells;
  c->stats.n_slow = n_slow;
  c->stats.n_fast_launches = c->last_launches;
  if (n_slow > b.slow_cap)
    return fail(c, MSVG_ENORESULT, "allocation-free search: the byte-MSV queue overflowed (" + std::to_string(n_slow) +
                                       " pairs, room for " + std::to_string(b.slow_cap) + "); use msvg_search");
  return MSVG_OK;
}

int msvg_results(msvg_ctx *c, float *usc, uint8_t *status, uint8_t *pass) {
  if (!c || c->last_db < 0 || !c->last) return fail(c, MSVG_EINVAL, "no results");
  Bufs &b = *c->last;
  const size_t cnt = (size_t) c->last_np * c->dbs[c->last_db]->n;
  if (usc) HIPCHK(c, hipMemcpy(usc, b.d_usc.p, cnt * sizeof(float), hipMemcpyDeviceToHost));
  if (status) HIPCHK(c, hipMemcpy(status, b.d_status.p, cnt, hipMemcpyDeviceToHost));
  if (pass) HIPCHK(c, hipMemcpy(pass, b.d_pass.p, cnt, hipMemcpyDeviceToHost));
  if (status) {
    for (size_t i = 0; i < cnt; ++i)
      if (status[i] == 255) return fail(c, MSVG_EDEVICE, "a pair was left unresolved by the device stages");
  }
  return MSVG_OK;
}

int msvg_search(msvg_ctx *c, const int32_t *ids, int32_t np, int32_t dbid, double F1, int mode, float *usc,
                uint8_t *status, uint8_t *pass) {
  int rc = msvg_search_device(c, ids, np, dbid, F1, mode);
  if (rc != MSVG_OK) return rc;
  std::vector<uint8_t> st;
  if (!status) { st.resize((size_t) np * c->dbs[dbid]->n); status = st.data(); }
  return msvg_results(c, usc, status, pass);
}

int msvg_last_stats(msvg_ctx *c, msvg_stats *out) {
  if (!c || !out) return MSVG_EINVAL;
  *out = c->stats;
  return MSVG_OK;
}

int msvg_set_tail_guard(msvg_ctx *c, int on) {
  if (!c) return MSVG_EINVAL;
  c->tail_guard = on != 0;
  c->last = nullptr;
  c->plans.clear();
  return MSVG_OK;
}

int msvg_set_tile(msvg_ctx *c, int32_t lo, int32_t hi, int32_t lanes, int32_t waves) {
  if (!c) return MSVG_EINVAL;
  if (lanes == 0) { c->tiles.clear(); c->last = nullptr; c->plans.clear(); return MSVG_OK; }
  if (lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64) return fail(c, MSVG_EINVAL, "lanes must be 8, 16, 32 or 64");
  if (waves < 0 || waves > 16) return fail(c, MSVG_EINVAL, "waves must be 0..16");
  c->tiles.push_back(Tile{lo, hi, lanes, waves});
  c->last = nullptr;
  c->plans.clear();
  return MSVG_OK;
}

int msvg_set_vit_tile(msvg_ctx *c, int32_t lo, int32_t hi, int32_t lanes, int32_t waves) {
  if (!c) return MSVG_EINVAL;
  if (lanes == 0) { c->vtiles.clear(); return MSVG_OK; }
  if (lanes != 16 && lanes != 32 && lanes != 64) return fail(c, MSVG_EINVAL, "lanes must be 16, 32 or 64");
  if (lanes == 64 && waves != 1 && waves != 2 && waves != 4 && waves != 8)
    return fail(c, MSVG_EINVAL, "waves must be 1, 2, 4 or 8");
  c->vtiles.push_back(Tile{lo, hi, lanes, waves});
  return MSVG_OK;
}

int msvg_set_vit_vote(msvg_ctx *c, int32_t mode) {
  if (!c) return MSVG_EINVAL;
  if (mode < -1 || mode > 1) return fail(c, MSVG_EINVAL, "mode must be -1, 0 or 1");
  c->vit_vote = mode;
  return MSVG_OK;
}

int msvg_set_vit_segments(msvg_ctx *c, int32_t on) {
  if (!c) return MSVG_EINVAL;
  if (on != 0 && on != 1) return fail(c, MSVG_EINVAL, "on must be 0 or 1");
  c->vit_segments = on;
  return MSVG_OK;
}

int msvg_profile_set_evparams(msvg_ctx *c, int32_t id, float mmu, float mlambda) {
  if (!c || !c->profiles.count(id)) return fail(c, MSVG_EINVAL, "unknown profile id");
  c->profiles[id]->mu = mmu;
  c->profiles[id]->lambda = mlambda;
  if (c->scanset) c->scanset->rows_ok = c->scanset->grows_ok = false;
  for (auto it = c->finals.begin(); it != c->finals.end();)
    it = (std::get<0>(it->first) == id) ? c->finals.erase(it) : std::next(it);
  c->last = nullptr;
  c->plans.clear();
  return MSVG_OK;
}

int msvg_calibrate_mu(msvg_ctx *c, int32_t id, double lambda, float *ret_mu) {
  if (!c || !ret_mu || !c->profiles.count(id)) return fail(c, MSVG_EINVAL, "unknown profile id");
  const int N = 200, L = 200;                       // EmN, EmL: evalues.c:67-68
  std::vector<uint8_t> res((size_t) N * L);
  msvh_calibration_targets(N, L, res.data());
  std::vector<int64_t> off(N);
  std::vector<int32_t> len(N, L);
  for (int i = 0; i < N; ++i) off[i] = (int64_t) i * L;
  int32_t db;
  int rc = msvg_db_add(c, res.data(), off.data(), len.data(), N, &db);
  if (rc != MSVG_OK) return rc;
  std::vector<float> usc(N);
  std::vector<uint8_t> st(N);
  rc = msvg_search(c, &id, 1, db, 0.02, MSVG_MODE_PLAIN, usc.data(), st.data(), nullptr);
  msvg_db_remove(c, db);
  if (rc != MSVG_OK) return rc;
  std::vector<int32_t> st32(st.begin(), st.end());
  const Profile &p = *c->profiles[id];
  *ret_mu = (float) msvh_fit_mu(usc.data(), st32.data(), N, L, p.base, p.scale, lambda);
  return MSVG_OK;
}

// ---------------------------------------------------------------------------
// ViterbiFilter stage
// ---------------------------------------------------------------------------

int msvg_profile_set_viterbi(msvg_ctx *c, int32_t pid, const msvg_vprofile *vp) {
  if (!c || !vp || !vp->rwv || !vp->twv) return fail(c, MSVG_EINVAL, "null argument");
  auto it = c->profiles.find(pid);
  if (it == c->profiles.end()) return fail(c, MSVG_EINVAL, "unknown profile id");
  Profile &p = *it->second;
  if (vp->M != p.M || vp->Kp != p.Kp) return fail(c, MSVG_EINVAL, "M or Kp differs from the profile's");
  if (!(vp->scale_w > 0.f)) return fail(c, MSVG_EINVAL, "scale_w must be positive");
  (void) hipSetDevice(c->device);
  auto v = std::make_unique<VitProfile>();
  v->rwv.assign(vp->rwv, vp->rwv + (size_t) p.Kp * p.M);
  v->twv.assign(vp->twv, vp->twv + (size_t) 8 * p.M);
  v->e_loop = vp->e_loop; v->e_move = vp->e_move; v->base_w = vp->base_w; v->ddbound_w = vp->ddbound_w;
  v->scale_w = vp->scale_w; v->nj = vp->nj; v->vmu = vp->vmu; v->vlambda = vp->vlambda;
  vit_shape(p.M, p.Kp, v->Q, v->W, v->G);
  // Rows vote on the lazy-F test unless it holds on most rows (then the
  // delete-path closure runs on every row, without the vote's work).
  v->vote = c->vit_vote >= 0 ? c->vit_vote
          : (msvh_vit_lazyf_rate(v->rwv.data(), v->twv.data(), p.M, p.Kp, v->e_loop, v->base_w, v->ddbound_w,
                                 msvh_vit_move(v->scale_w, v->nj, 350), 128) < 0.75 ? 1 : 0);
  for (const Tile &t : c->vtiles) {
    if (p.M < t.lo || p.M > t.hi) continue;
    v->G = t.lanes;
    v->W = t.lanes == 64 ? std::max(1, t.waves) : 1;
    v->Q = (p.M + 2 * vit_lanes(*v) - 1) / (2 * vit_lanes(*v));
    if (v->Q > (v->G == 64 && v->W == 1 ? vit_band64_qmax(p.Kp) : 16))
      return fail(c, MSVG_EINVAL, "Viterbi tile: too many registers per lane");
  }
  const size_t per = (size_t) v->Q * vit_lanes(*v);
  std::vector<uint32_t> h(8 * per);
  for (int t = 0; t < 8; ++t)
    vit_fill(h, t * per, v->Q, vit_lanes(*v), p.M, [&](int k) { return (int) v->twv[(size_t) t * p.M + k - 1]; });
  HIPCHK(c, v->d_ttab.ensure(h.size() * 4));
  HIPCHK(c, hipMemcpy(v->d_ttab.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  if (v->W > 1) {
    // Column segments for vit_seg: as few as 14 registers per lane (1792
    // positions) allow, of equal width (a multiple of 128 positions; the
    // last one takes the rest), each with its transitions at their model
    // positions.
    const int nseg = (p.M + 1791) / 1792;
    v->seg_span = 128 * (((p.M + nseg - 1) / nseg + 127) / 128);
    for (int k0 = 0; k0 < p.M; k0 += v->seg_span) {
      const int Ms = std::min(v->seg_span, p.M - k0), Qs = (Ms + 127) / 128;
      std::vector<uint32_t> hs((size_t) 8 * Qs * 64);
      for (int t = 0; t < 8; ++t)
        vit_fill(hs, (size_t) t * Qs * 64, Qs, 64, Ms,
                 [&](int k) { return (int) v->twv[(size_t) t * p.M + k0 + k - 1]; });
      auto mem = std::make_unique<DevMem>();
      HIPCHK(c, mem->ensure(hs.size() * 4));
      HIPCHK(c, hipMemcpy(mem->p, hs.data(), hs.size() * 4, hipMemcpyHostToDevice));
      v->d_seg_ttab.push_back(std::move(mem));
    }
  }
  p.vit = std::move(v);
  for (auto &kv : c->tables)                       // emission tables of the old fields
    if (kv.first.first == pid) { kv.second->vemit.reset(); kv.second->vseg_emit.clear(); }
  return MSVG_OK;
}

}  // extern "C"

namespace {

// One profile's share of a ViterbiFilter search: its items (longest target
// first), their N/C/J move costs, where they sit in the device item arrays,
// and, for the multi-wave shapes, the launch.
struct VitJob {
  int pid = 0;
  int64_t n = 0, base = 0;           // items [base, base + n) of the device arrays
  std::vector<int64_t> order;        // item q -> caller position
  std::vector<int64_t> items;        // item q -> sorted target index
  std::vector<int16_t> mv;
  std::vector<int32_t> uoff;         // multi-wave shapes: workgroup k runs items [uoff[k], uoff[k+1])
                                     // (column segments: wave slot k)
  int64_t ubase = 0;                 // where uoff sits in the device index array
  int64_t cells = 0;
  VitArgs a{};
  // A multi-wave shape run as column segments (vit_seg): the launches'
  // workgroups, each item's first carry row (from the job's own start) and
  // the job's carry rows, and where they start in the device carry arrays.
  bool seg = false;
  int seg_grid = 0;
  std::vector<int64_t> coff;
  int64_t crows = 0, cbase = 0;
};

// One vit_band launch: the jobs of one (G, Q, vote) class.  Workgroup b
// works on job wg_prof[b]'s profile; its waves run the units the plan lists
// for them (unit = 64/G consecutive items of one job).
struct BandLaunch {
  int G = 16, Q = 1, vote = 1, grid = 0;
  std::vector<int> jobs;             // indices into the search's jobs
  std::vector<int32_t> wg_prof, uoff, ustart;
  int64_t ibase = 0;                 // wg_prof, uoff, ustart in the device index array, in that order
  int64_t pbase = 0;                 // its VitProfDev entries in the device profile array
  VitArgs a{};
};

// The checks and device tables of one job (on the calling thread: they may
// build and upload the profile's emission table for the database).
int vit_check(msvg_ctx *c, int32_t pid, int32_t dbid, const int64_t *targets, int64_t n, VitJob &job) {
  if (n < 0 || (n > 0 && !targets)) return fail(c, MSVG_EINVAL, "null argument");
  auto pit = c->profiles.find(pid);
  if (pit == c->profiles.end()) return fail(c, MSVG_EINVAL, "unknown profile id");
  Profile &p = *pit->second;
  if (!p.vit) return fail(c, MSVG_EINVAL, "profile has no ViterbiFilter fields (msvg_profile_set_viterbi)");
  auto dit = c->dbs.find(dbid);
  if (dit == c->dbs.end()) return fail(c, MSVG_EINVAL, "unknown database id");
  Db &d = *dit->second;
  if (d.maxcode >= p.Kp) return fail(c, MSVG_EINVAL, "database holds residue codes outside the profile's alphabet");
  for (int64_t i = 0; i < n; ++i)
    if (targets[i] < 0 || targets[i] >= d.n) return fail(c, MSVG_EINVAL, "target index out of range");
  job.pid = pid;
  job.n = n;
  if (n == 0) return MSVG_OK;
  VitProfile &v = *p.vit;
  const uint32_t *etab;
  int rc = vit_emission_table(c, pid, dbid, &etab);
  if (rc != MSVG_OK) return rc;
  job.seg = v.W > 1 && c->vit_segments && v.seg_span > 0;
  if (job.seg && (rc = vit_seg_emission_tables(c, pid, dbid)) != MSVG_OK) return rc;
  VitArgs &a = job.a;
  a.etab = etab;
  a.ttab = v.d_ttab.as<uint32_t>();
  a.db.res = d.d_res.as<uint8_t>(); a.db.off = d.d_off.as<int64_t>(); a.db.len = d.d_len.as<int32_t>();
  a.db.tjb = d.d_tjb.as<uint8_t>(); a.db.n = d.n; a.db.R = d.R;
  a.n = (int32_t) n;
  a.e_loop = v.e_loop; a.e_move = v.e_move; a.base_w = v.base_w; a.ddbound_w = v.ddbound_w;
  return MSVG_OK;
}

// The host side of one checked job (no device calls, so jobs run on host
// threads side by side): items longest target first -- units (one wave's
// targets) are consecutive items, so a unit's rows are its first target's
// length --, their move costs (one per distinct length) and cells, and for
// the multi-wave shapes the lists of the job's own launch.
void vit_fill(const VitProfile &v, int M, const Db &d, int slots, const int64_t *targets, VitJob &job) {
  const int64_t n = job.n;
  if (n == 0) return;
  std::vector<std::pair<int32_t, int64_t>> key((size_t) n);       // (-length, caller position)
  for (int64_t q = 0; q < n; ++q) key[(size_t) q] = {-d.len_sorted[(size_t) d.sorted_of[(size_t) targets[q]]], q};
  std::sort(key.begin(), key.end());
  job.order.resize((size_t) n);
  for (int64_t q = 0; q < n; ++q) job.order[(size_t) q] = key[(size_t) q].second;
  if (v.W > 1) {
    // Multi-wave shapes, one launch per profile: items (one per workgroup at
    // a time) go greedily to the launch's workgroups, longest first, each to
    // the workgroup with the fewest rows so far, any number per workgroup;
    // workgroup k runs items [uoff[k], uoff[k+1]) of the reordered list.
    std::vector<std::vector<int64_t>> lists((size_t) slots);
    using Slot = std::pair<int64_t, int>;                       // (rows so far, slot)
    std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
    for (int k = 0; k < slots; ++k) heap.push(Slot{0, k});
    for (int64_t u = 0; u < n; ++u) {
      Slot sl = heap.top();
      heap.pop();
      lists[(size_t) sl.second].push_back(u);
      sl.first += -key[(size_t) u].first;
      heap.push(sl);
    }
    job.uoff.assign((size_t) slots + 1, 0);
    std::vector<int64_t> dealt;
    dealt.reserve((size_t) n);
    for (int k = 0; k < slots; ++k) {
      job.uoff[(size_t) k + 1] = job.uoff[(size_t) k] + (int32_t) lists[(size_t) k].size();
      for (const int64_t u : lists[(size_t) k]) dealt.push_back(job.order[(size_t) u]);
    }
    job.order.swap(dealt);
  }
  job.items.resize((size_t) n);
  job.mv.resize((size_t) n);
  std::vector<int32_t> mv_of((size_t) d.Ls.size(), INT32_MIN);   // move cost per distinct length
  for (int64_t q = 0; q < n; ++q) {
    const int64_t s = d.sorted_of[(size_t) targets[job.order[(size_t) q]]];
    job.items[(size_t) q] = s;
    int32_t &m = mv_of[(size_t) d.lidx_sorted[(size_t) s]];
    if (m == INT32_MIN) m = msvh_vit_move(v.scale_w, v.nj, d.len_sorted[(size_t) s]);
    job.mv[(size_t) q] = (int16_t) m;
    job.cells += (int64_t) M * d.len_sorted[(size_t) s];
  }
  if (job.seg) {
    job.coff.resize((size_t) n);
    int64_t r = 0;
    for (int64_t q = 0; q < n; ++q) {
      job.coff[(size_t) q] = r;
      r += (d.len_sorted[(size_t) job.items[(size_t) q]] + 15) & ~15;
    }
    job.crows = r;
  }
}

// vit_band launches for the W = 1 jobs: jobs of one (G, Q, vote) class
// share a launch, so that the row chains of their longest targets run beside
// each other and beside the other jobs' work instead of ending a launch
// each.  Workgroups (each works on one job) are given out one at a time:
// first one to every job, then each to the job with the most rows per wave
// while it has units left for the workgroup's waves, up to what the device
// holds at once (a launch that leaves SIMDs with fewer waves runs each of its
// waves faster but issues less: C2 measured 5.5 ms on 449 of 512 workgroups
// against 5.1 ms on all).  Each job's units then go greedily to its own waves:
// longest first, each to the wave with the fewest rows so far.  A class with
// more jobs than the device holds workgroups is split.  Positions (job.base)
// must be assigned.
int vit_plan_band(msvg_ctx *c, const std::vector<VitJob> &jobs, const Db &d, std::vector<BandLaunch> &out) {
  std::map<std::tuple<int, int, int>, std::vector<int>> classes;
  for (size_t i = 0; i < jobs.size(); ++i) {
    if (jobs[i].n == 0) continue;
    const VitProfile &v = *c->profiles[jobs[i].pid]->vit;
    if (v.W == 1) classes[std::make_tuple(v.G, v.Q, v.vote)].push_back((int) i);
  }
  for (auto &kv : classes) {
    const int G = std::get<0>(kv.first), Q = std::get<1>(kv.first), vote = std::get<2>(kv.first);
    const int nw = msvk_vit_band_block(G) / 64, NG = 64 / G;
    const int cap = msvk_vit_band_cap(G, Q, vote, d.R, c->ncu);
    if (cap < 1) return fail(c, MSVG_EINVAL, "ViterbiFilter shape has no launch");
    const std::vector<int> &all = kv.second;
    for (size_t first = 0; first < all.size(); first += (size_t) cap) {
      BandLaunch L;
      L.G = G; L.Q = Q; L.vote = vote;
      L.jobs.assign(all.begin() + (long) first, all.begin() + (long) std::min(all.size(), first + (size_t) cap));
      const int nj = (int) L.jobs.size();
      auto unit_rows = [&](const VitJob &j, int64_t u) { return (int64_t) d.len_sorted[j.items[(size_t) (u * NG)]]; };
      std::vector<int64_t> nunit((size_t) nj), rows((size_t) nj, 0), wg((size_t) nj, 1);
      int64_t spare = cap - nj;
      for (int q = 0; q < nj; ++q) {
        const VitJob &j = jobs[(size_t) L.jobs[(size_t) q]];
        nunit[(size_t) q] = (j.n + NG - 1) / NG;
        for (int64_t u = 0; u < nunit[(size_t) q]; ++u) rows[(size_t) q] += unit_rows(j, u);
      }
      auto load = [&](int q) { return (double) rows[(size_t) q] / (double) (wg[(size_t) q] * nw); };
      auto gains = [&](int q) { return wg[(size_t) q] * nw < nunit[(size_t) q]; };
      std::priority_queue<std::pair<double, int>> heap;
      for (int q = 0; q < nj; ++q) if (gains(q)) heap.push({load(q), q});
      while (spare > 0 && !heap.empty()) {
        const int q = heap.top().second;
        heap.pop();
        ++wg[(size_t) q];
        --spare;
        if (gains(q)) heap.push({load(q), q});
      }
      int64_t nwg = 0;
      for (int q = 0; q < nj; ++q) nwg += wg[(size_t) q];
      L.grid = (int) nwg;
      L.wg_prof.resize((size_t) nwg);
      L.uoff.assign((size_t) nwg * nw + 1, 0);
      int64_t b0 = 0;
      for (int q = 0; q < nj; ++q) {
        const VitJob &j = jobs[(size_t) L.jobs[(size_t) q]];
        const int S = (int) (wg[(size_t) q] * nw);
        std::vector<std::vector<int64_t>> lists((size_t) S);
        using Slot = std::pair<int64_t, int>;                   // (rows so far, wave)
        std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap2;
        for (int k = 0; k < S; ++k) heap2.push(Slot{0, k});
        for (int64_t u = 0; u < nunit[(size_t) q]; ++u) {
          Slot sl = heap2.top();
          heap2.pop();
          lists[(size_t) sl.second].push_back(u);
          sl.first += unit_rows(j, u);
          heap2.push(sl);
        }
        for (int64_t b = b0; b < b0 + wg[(size_t) q]; ++b) L.wg_prof[(size_t) b] = q;
        for (int k = 0; k < S; ++k) {
          const int64_t slot = b0 * nw + k;
          L.uoff[(size_t) slot + 1] = L.uoff[(size_t) slot] + (int32_t) lists[(size_t) k].size();
          for (const int64_t u : lists[(size_t) k]) L.ustart.push_back((int32_t) (j.base + u * NG));
        }
        b0 += wg[(size_t) q];
      }
      out.push_back(std::move(L));
    }
  }
  return MSVG_OK;
}

// Device arrays of a search, uploaded on st: items and move costs of all
// jobs, the multi-wave jobs' lists, and the vit_band launches' profiles,
// workgroup profiles, lists and unit starts.
int vit_upload(msvg_ctx *c, std::vector<VitJob> &jobs, std::vector<BandLaunch> &bands, hipStream_t st,
               std::vector<int64_t> &hitems, std::vector<int16_t> &hmv, std::vector<int32_t> &hidx,
               std::vector<VitProfDev> &hprof) {
  int64_t tot = 0;
  for (const VitJob &j : jobs) tot += j.n;
  hidx.clear();
  for (VitJob &j : jobs) { j.ubase = (int64_t) hidx.size(); hidx.insert(hidx.end(), j.uoff.begin(), j.uoff.end()); }
  hprof.clear();
  for (BandLaunch &L : bands) {
    L.ibase = (int64_t) hidx.size();
    hidx.insert(hidx.end(), L.wg_prof.begin(), L.wg_prof.end());
    hidx.insert(hidx.end(), L.uoff.begin(), L.uoff.end());
    hidx.insert(hidx.end(), L.ustart.begin(), L.ustart.end());
    L.pbase = (int64_t) hprof.size();
    for (const int ji : L.jobs) {
      const VitJob &j = jobs[(size_t) ji];
      VitProfDev pd{};
      pd.etab = j.a.etab; pd.ttab = j.a.ttab;
      pd.e_loop = j.a.e_loop; pd.e_move = j.a.e_move; pd.base_w = j.a.base_w; pd.ddbound_w = j.a.ddbound_w;
      pd.end = (int32_t) (j.base + j.n);
      hprof.push_back(pd);
    }
  }
  if (tot > INT32_MAX) return fail(c, MSVG_EINVAL, "too many ViterbiFilter pairs in one call");
  // Column-segmented jobs: carry rows, each item's first carry row, partial maxima.
  int64_t crows = 0;
  for (VitJob &j : jobs)
    if (j.seg) { j.cbase = crows; crows += j.crows; }
  if (crows > 0) {
    std::vector<int64_t> hcoff((size_t) tot, 0);
    for (const VitJob &j : jobs)
      if (j.seg)
        for (int64_t q = 0; q < j.n; ++q) hcoff[(size_t) (j.base + q)] = j.cbase + j.coff[(size_t) q];
    HIPCHK(c, c->d_vcoff.ensure((size_t) tot * sizeof(int64_t)));
    HIPCHK(c, c->d_vpartial.ensure((size_t) tot * sizeof(int32_t)));
    HIPCHK(c, c->d_vcarry.ensure((size_t) crows * 4 * sizeof(int16_t)));
    HIPCHK(c, hipMemcpyAsync(c->d_vcoff.p, hcoff.data(), (size_t) tot * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipStreamSynchronize(st));               // hcoff ends with this block
  }
  c->vcrows = crows;
  HIPCHK(c, c->d_vuoff.ensure(std::max<size_t>(hidx.size(), 1) * sizeof(int32_t)));
  HIPCHK(c, c->d_vprofs.ensure(std::max<size_t>(hprof.size(), 1) * sizeof(VitProfDev)));
  HIPCHK(c, c->d_vitems.ensure((size_t) std::max<int64_t>(tot, 1) * sizeof(int64_t)));
  HIPCHK(c, c->d_vmove.ensure((size_t) std::max<int64_t>(tot, 1) * sizeof(int16_t)));
  HIPCHK(c, c->d_vout.ensure((size_t) std::max<int64_t>(tot, 1) * sizeof(int32_t)));
  hitems.resize((size_t) tot);
  hmv.resize((size_t) tot);
  for (const VitJob &j : jobs) {
    std::copy(j.items.begin(), j.items.end(), hitems.begin() + j.base);
    std::copy(j.mv.begin(), j.mv.end(), hmv.begin() + j.base);
  }
  if (!hidx.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_vuoff.p, hidx.data(), hidx.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (!hprof.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_vprofs.p, hprof.data(), hprof.size() * sizeof(VitProfDev), hipMemcpyHostToDevice, st));
  if (tot == 0) return MSVG_OK;
  HIPCHK(c, hipMemcpyAsync(c->d_vitems.p, hitems.data(), (size_t) tot * sizeof(int64_t), hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(c->d_vmove.p, hmv.data(), (size_t) tot * sizeof(int16_t), hipMemcpyHostToDevice, st));
  return MSVG_OK;
}

int vit_launch_band(msvg_ctx *c, BandLaunch &L, const VitJob &any, hipStream_t st) {
  VitArgs &a = L.a;
  a = VitArgs{};
  a.db = any.a.db;
  a.items = c->d_vitems.as<int64_t>();
  a.move = c->d_vmove.as<int16_t>();
  a.out = c->d_vout.as<int32_t>();
  const int32_t *idx = c->d_vuoff.as<int32_t>() + L.ibase;
  a.wg_prof = idx;
  a.uoff = idx + L.wg_prof.size();
  a.ustart = idx + L.wg_prof.size() + L.uoff.size();
  a.profs = c->d_vprofs.as<VitProfDev>() + L.pbase;
  HIPCHK(c, msvk_launch_vit_band(L.G, L.Q, L.vote, &a, L.grid, st));
  return MSVG_OK;
}

int vit_launch_multiwave(msvg_ctx *c, VitJob &job, hipStream_t st) {
  const VitProfile &v = *c->profiles[job.pid]->vit;
  job.a.items = c->d_vitems.as<int64_t>() + job.base;
  job.a.move = c->d_vmove.as<int16_t>() + job.base;
  job.a.out = c->d_vout.as<int32_t>() + job.base;
  job.a.uoff = c->d_vuoff.as<int32_t>() + job.ubase;
  HIPCHK(c, msvk_launch_vit(v.Q, v.W, &job.a, c->ncu, st));
  return MSVG_OK;
}

// A column-segmented job: its segments in order on st (each reads the
// carries the one before wrote).
int vit_launch_segments(msvg_ctx *c, VitJob &job, int dbid, hipStream_t st) {
  const Profile &p = *c->profiles[job.pid];
  const VitProfile &v = *p.vit;
  const Tables &t = tables_for(c, job.pid, dbid);
  VitSegArgs a{};
  a.db = job.a.db;
  a.e_loop = v.e_loop; a.e_move = v.e_move; a.base_w = v.base_w; a.ddbound_w = v.ddbound_w;
  a.items = c->d_vitems.as<int64_t>() + job.base;
  a.move = c->d_vmove.as<int16_t>() + job.base;
  a.coff = c->d_vcoff.as<int64_t>() + job.base;
  int16_t *cb = c->d_vcarry.as<int16_t>();
  a.cM = cb; a.cI = cb + c->vcrows; a.cD = cb + 2 * c->vcrows; a.cE = cb + 3 * c->vcrows;
  a.partial = c->d_vpartial.as<int32_t>() + job.base;
  a.out = c->d_vout.as<int32_t>() + job.base;
  a.uoff = c->d_vuoff.as<int32_t>() + job.ubase;
  size_t si = 0;
  for (int k0 = 0; k0 < p.M; k0 += v.seg_span, ++si) {
    if (si >= t.vseg_emit.size() || si >= v.d_seg_ttab.size()) return fail(c, MSVG_EINVAL, "segment tables missing");
    const int Qs = (std::min(v.seg_span, p.M - k0) + 127) / 128;
    a.etab = t.vseg_emit[si]->as<uint32_t>();
    a.ttab = v.d_seg_ttab[si]->as<uint32_t>();
    a.first = (k0 == 0);
    a.last = (k0 + v.seg_span >= p.M);
    HIPCHK(c, msvk_launch_vit_seg(Qs, &a, job.seg_grid, st));
  }
  return MSVG_OK;
}

// Workgroups of a column-segmented job's launches: what the device holds of
// its widest segment, and no more than its items fill (eight per workgroup).
int vit_seg_grid(msvg_ctx *c, const VitProfile &v, int M, int R, int64_t n) {
  int cap = INT32_MAX;
  for (int k0 = 0; k0 < M; k0 += v.seg_span) cap = std::min(cap, msvk_vit_seg_cap((std::min(v.seg_span, M - k0) + 127) / 128, R, c->ncu));
  if (cap < 1) return 0;
  return (int) std::max<int64_t>(1, std::min<int64_t>(cap, (n + 7) / 8));
}

// Scores of job j from the raw results (host expression of vitfilter.c:240-248).
void vit_finish(msvg_ctx *c, const VitJob &j, const int32_t *out, float *sc, uint8_t *status) {
  const VitProfile &v = *c->profiles[j.pid]->vit;
  for (int64_t q = 0; q < j.n; ++q) {
    const int64_t i = j.order[(size_t) q];
    const int32_t r = out[j.base + q];
    if (r == VIT_DEFER) continue;                              // scored again by the exact kernel
    if ((r >> 16) == MSVG_ERANGE) {
      status[i] = MSVG_ERANGE;
      sc[i] = INFINITY;                                        // vitfilter.c:177
    } else {
      status[i] = MSVG_OK;
      sc[i] = msvh_vit_score((int16_t) (r & 0xffff), j.mv[(size_t) q], v.base_w, v.scale_w);
    }
  }
}

}  // namespace

extern "C" {

int msvg_vit_search(msvg_ctx *c, int32_t pid, int32_t dbid, const int64_t *targets, int64_t n, float *sc,
                    uint8_t *status) {
  if (!c) return MSVG_EINVAL;
  if (n > 0 && (!sc || !status)) return fail(c, MSVG_EINVAL, "null argument");
  const int64_t offs[2] = {0, n};
  return msvg_vit_search_many(c, 1, &pid, dbid, offs, targets, sc, status);
}

int msvg_vit_search_many(msvg_ctx *c, int32_t np, const int32_t *pids, int32_t dbid, const int64_t *offsets,
                         const int64_t *targets, float *sc, uint8_t *status) {
  if (!c) return MSVG_EINVAL;
  if (np < 0 || (np > 0 && (!pids || !offsets))) return fail(c, MSVG_EINVAL, "null argument");
  for (int32_t i = 0; i < np; ++i)
    if (offsets[i + 1] < offsets[i] || offsets[i] < 0) return fail(c, MSVG_EINVAL, "offsets must not decrease");
  const int64_t total = np > 0 ? offsets[np] - offsets[0] : 0;
  if (total > 0 && (!targets || !sc || !status)) return fail(c, MSVG_EINVAL, "null argument");
  c->stats = msvg_stats{};
  (void) hipSetDevice(c->device);
  std::vector<VitJob> jobs((size_t) np);
  for (int32_t i = 0; i < np; ++i) {
    const int rc = vit_check(c, pids[i], dbid, targets + offsets[i], offsets[i + 1] - offsets[i], jobs[(size_t) i]);
    if (rc != MSVG_OK) return rc;
  }
  if (total == 0) return MSVG_OK;
  {
    // The jobs' host preparation on up to 16 threads, jobs claimed one at a time.
    const Db &d = *c->dbs[dbid];
    std::vector<int> slots((size_t) np, 0);
    for (int32_t i = 0; i < np; ++i) {
      const VitProfile &v = *c->profiles[pids[i]]->vit;
      if (jobs[(size_t) i].n > 0 && v.W > 1) {
        if (jobs[(size_t) i].seg) {
          jobs[(size_t) i].seg_grid = vit_seg_grid(c, v, c->profiles[pids[i]]->M, d.R, jobs[(size_t) i].n);
          slots[(size_t) i] = 8 * jobs[(size_t) i].seg_grid;
        } else {
          slots[(size_t) i] = msvk_vit_slots(v.Q, v.W, &jobs[(size_t) i].a, c->ncu);
        }
        if (slots[(size_t) i] < 1) return fail(c, MSVG_EINVAL, "ViterbiFilter shape has no launch");
      }
    }
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int i; (i = next.fetch_add(1)) < np;) {
        const Profile &p = *c->profiles[pids[i]];
        vit_fill(*p.vit, p.M, d, slots[(size_t) i], targets + offsets[i], jobs[(size_t) i]);
      }
    };
    const int nth = std::max(1, std::min({16, (int) np, (int) std::max(1u, std::thread::hardware_concurrency()),
                                          (int) (total / 20000) + 1}));
    std::vector<std::thread> th;
    try {
      for (int t = 1; t < nth; ++t) th.emplace_back(work);
    } catch (...) {                    // a thread that cannot start: the calling thread takes its jobs
    }
    work();
    for (auto &x : th) x.join();
  }
  int64_t pos = 0;
  for (VitJob &j : jobs) { j.base = pos; pos += j.n; }
  std::vector<BandLaunch> bands;
  int rc = vit_plan_band(c, jobs, *c->dbs[dbid], bands);
  if (rc != MSVG_OK) return rc;
  hipStream_t st = c->stream;
  std::vector<int64_t> hitems;
  std::vector<int16_t> hmv;
  std::vector<int32_t> hidx;
  std::vector<VitProfDev> hprof;
  if ((rc = vit_upload(c, jobs, bands, st, hitems, hmv, hidx, hprof)) != MSVG_OK) return rc;
  // The vit_band launches (one per shape class), then the multi-wave
  // profiles' launches, round-robin over the context's ViterbiFilter streams
  // so that the launches' tails (each lasts at least as long as its longest
  // target's row chain) overlap.
  const VitJob *any = nullptr;
  std::vector<int> multi;
  for (size_t i = 0; i < jobs.size(); ++i) {
    if (jobs[i].n == 0) continue;
    any = &jobs[i];
    if (c->profiles[jobs[i].pid]->vit->W > 1) multi.push_back((int) i);
  }
  const int nlaunch = (int) (bands.size() + multi.size());
  const int nstreams = nlaunch > 1 ? kVitStreams : 1;
  if (nstreams > 1 && !c->vstream[0]) {
    for (int k = 0; k < kVitStreams; ++k) {
      HIPCHK(c, hipStreamCreateWithFlags(&c->vstream[k], hipStreamNonBlocking));
      HIPCHK(c, hipEventCreateWithFlags(&c->vdone[k], hipEventDisableTiming));
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[0], st));
  if (nstreams > 1) {
    HIPCHK(c, hipEventRecord(c->fork, st));
    for (int k = 0; k < nstreams; ++k) HIPCHK(c, hipStreamWaitEvent(c->vstream[k], c->fork, 0));
  }
  int k = 0;
  for (BandLaunch &L : bands)
    if ((rc = vit_launch_band(c, L, *any, nstreams > 1 ? c->vstream[k++ % nstreams] : st)) != MSVG_OK) return rc;
  for (const int i : multi) {
    hipStream_t ms = nstreams > 1 ? c->vstream[k++ % nstreams] : st;
    rc = jobs[(size_t) i].seg ? vit_launch_segments(c, jobs[(size_t) i], dbid, ms)
                              : vit_launch_multiwave(c, jobs[(size_t) i], ms);
    if (rc != MSVG_OK) return rc;
  }
  if (nstreams > 1) {
    for (int q = 0; q < nstreams; ++q) {
      HIPCHK(c, hipEventRecord(c->vdone[q], c->vstream[q]));
      HIPCHK(c, hipStreamWaitEvent(st, c->vdone[q], 0));
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[1], st));
  std::vector<int32_t> out((size_t) total);
  HIPCHK(c, hipMemcpyAsync(out.data(), c->d_vout.p, (size_t) total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  int64_t cells = 0;
  for (int32_t i = 0; i < np; ++i) {
    vit_finish(c, jobs[(size_t) i], out.data(), sc + offsets[i], status + offsets[i]);
    cells += jobs[(size_t) i].cells;
  }
  // Items of column-segmented jobs some row of which could have changed B
  // (VIT_DEFER): scored again by the exact multi-wave kernel.
  std::vector<int32_t> dpid;
  std::vector<int64_t> doffs{0}, dtarg, dpos;
  for (int32_t i = 0; i < np; ++i) {
    const VitJob &j = jobs[(size_t) i];
    if (!j.seg) continue;
    for (int64_t q = 0; q < j.n; ++q)
      if (out[(size_t) (j.base + q)] == VIT_DEFER) {
        const int64_t at = offsets[i] + j.order[(size_t) q];
        dtarg.push_back(targets[at]);
        dpos.push_back(at);
      }
    if ((int64_t) dtarg.size() > doffs.back()) { dpid.push_back(pids[i]); doffs.push_back((int64_t) dtarg.size()); }
  }
  if (!dtarg.empty()) {
    std::vector<float> dsc(dtarg.size());
    std::vector<uint8_t> dst(dtarg.size());
    const int keep = c->vit_segments;
    c->vit_segments = 0;
    rc = msvg_vit_search_many(c, (int32_t) dpid.size(), dpid.data(), dbid, doffs.data(), dtarg.data(), dsc.data(),
                              dst.data());
    c->vit_segments = keep;
    if (rc != MSVG_OK) return rc;
    for (size_t u = 0; u < dpos.size(); ++u) { sc[dpos[u]] = dsc[u]; status[dpos[u]] = dst[u]; }
    ms += (float) c->stats.ms_fast;
  }
  c->stats = msvg_stats{};
  c->stats.ms_fast = ms;
  c->stats.ms_total = ms;
  c->stats.cells = cells;
  c->stats.n_slow = (int64_t) dtarg.size();
  c->stats.n_fast_launches = nlaunch;
  return MSVG_OK;
}

int msvg_ViterbiFilter(msvg_ctx *c, int32_t pid, const uint8_t *dsq, int L, float *ret_sc) {
  if (!c || !dsq || !ret_sc || L < 0) return fail(c, MSVG_EINVAL, "null argument");
  if (!c->profiles.count(pid)) return fail(c, MSVG_EINVAL, "unknown profile id");
  if (c->scratch_db >= 0) msvg_db_remove(c, c->scratch_db);
  c->scratch_db = -1;
  const int64_t off = 0;
  int32_t id;
  int rc = msvg_db_add(c, dsq + 1, &off, &L, 1, &id);
  if (rc != MSVG_OK) return rc;
  c->scratch_db = id;
  const int64_t t = 0;
  uint8_t st;
  rc = msvg_vit_search(c, pid, id, &t, 1, ret_sc, &st);
  if (rc != MSVG_OK) return rc;
  return st == MSVG_ERANGE ? MSVG_ERANGE : MSVG_OK;
}

int msvg_calibrate_vmu(msvg_ctx *c, int32_t id, double lambda, float *ret_vmu) {
  if (!c || !ret_vmu || !c->profiles.count(id)) return fail(c, MSVG_EINVAL, "unknown profile id");
  if (!c->profiles[id]->vit) return fail(c, MSVG_EINVAL, "profile has no ViterbiFilter fields");
  const int N = 200, L = 200;                       // EvN, EvL: evalues.c:69-70
  std::vector<uint8_t> res((size_t) N * L);
  msvh_vit_calibration_targets(N, L, res.data());
  std::vector<int64_t> off(N), tg(N);
  std::vector<int32_t> len(N, L);
  for (int i = 0; i < N; ++i) { off[i] = (int64_t) i * L; tg[i] = i; }
  int32_t db;
  int rc = msvg_db_add(c, res.data(), off.data(), len.data(), N, &db);
  if (rc != MSVG_OK) return rc;
  std::vector<float> sc(N);
  std::vector<uint8_t> st(N);
  rc = msvg_vit_search(c, id, db, tg.data(), N, sc.data(), st.data());
  msvg_db_remove(c, db);
  if (rc != MSVG_OK) return rc;
  std::vector<int32_t> st32(st.begin(), st.end());
  const VitProfile &v = *c->profiles[id]->vit;
  *ret_vmu = (float) msvh_fit_vmu(sc.data(), st32.data(), N, L, v.base_w, v.scale_w, lambda);
  return MSVG_OK;
}

int msvg_tile_of(msvg_ctx *c, int32_t M, int32_t dbid, int32_t *lanes, int32_t *waves, int32_t *segmented) {
  if (!c || !lanes || !waves || !segmented || M < 1) return MSVG_EINVAL;
  auto it = c->dbs.find(dbid);
  if (it == c->dbs.end()) return fail(c, MSVG_EINVAL, "unknown database id");
  int G = 64, w = 0;
  choose_tile(c, M, *it->second, G, w);
  const int Q = qfor(M, 2 * G);
  const int seg = (Q > kQMax || fast_lds(it->second->R, Q) > (size_t) kLdsBytes) ? 1 : 0;
  *lanes = seg ? 64 : G;
  *waves = w;
  *segmented = seg;
  return MSVG_OK;
}

int msvg_probe_lds(msvg_ctx *c, double *ret_bytes_per_s) {
  if (!c || !ret_bytes_per_s) return MSVG_EINVAL;
  (void) hipSetDevice(c->device);
  const int grid = c->ncu, iters = 4096;
  DevMem out;
  HIPCHK(c, out.ensure((size_t) grid * 1024 * 16));
  HIPCHK(c, msvk_lds_probe(out.p, grid, iters, c->stream));          // warm-up
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  HIPCHK(c, msvk_lds_probe(out.p, grid, iters, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  *ret_bytes_per_s = (double) grid * 1024 * iters * 8 * 16 / (ms / 1e3);
  return MSVG_OK;
}

int msvg_probe_forms(msvg_ctx *c, int32_t L, double *ms_f16, double *ms_u16, int64_t *cells, int32_t *agree) {
  if (!c || !ms_f16 || !ms_u16 || !cells || !agree || L < 16 || L % 16 != 0 || L > (1 << 20)) return MSVG_EINVAL;
  (void) hipSetDevice(c->device);
  constexpr int kRows = 21, kQB = 4, kWaves = 16;
  const int grid = c->ncu, nw = grid * kWaves;
  const int tab_u4 = kRows * kQB * 64;
  const int bias = 20;
  // Costs per (row, register, slot, lane) and residues from a fixed LCG;
  // every residue code is a table row (< kRows).
  uint32_t x = 12345u;
  auto next = [&x]() { x = x * 1664525u + 1013904223u; return x >> 8; };
  std::vector<uint32_t> t16((size_t) tab_u4 * 4), tu((size_t) tab_u4 * 4);
  for (size_t i = 0; i < t16.size(); ++i) {
    const uint32_t c0 = next() % 256u, c1 = next() % 256u;
    tu[i] = c0 | (c1 << 16);
    const _Float16 h0 = (_Float16) ((float) (bias - (int) c0) / 2048.0f);
    const _Float16 h1 = (_Float16) ((float) (bias - (int) c1) / 2048.0f);
    t16[i] = (uint32_t) __builtin_bit_cast(uint16_t, h0) | ((uint32_t) __builtin_bit_cast(uint16_t, h1) << 16);
  }
  std::vector<uint8_t> res((size_t) nw * L);
  for (auto &b : res) b = (uint8_t) (next() % (uint32_t) kRows);
  DevMem d_res, d_t16, d_tu, d_out0, d_out1;
  HIPCHK(c, d_res.ensure(res.size()));
  HIPCHK(c, d_t16.ensure(t16.size() * 4));
  HIPCHK(c, d_tu.ensure(tu.size() * 4));
  HIPCHK(c, d_out0.ensure((size_t) nw * 4));
  HIPCHK(c, d_out1.ensure((size_t) nw * 4));
  HIPCHK(c, hipMemcpy(d_res.p, res.data(), res.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_t16.p, t16.data(), t16.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_tu.p, tu.data(), tu.size() * 4, hipMemcpyHostToDevice));
  const uint32_t bias2 = (uint32_t) bias | ((uint32_t) bias << 16);
  double ms[2];
  for (int form = 0; form < 2; ++form) {
    const void *tab = form ? d_tu.p : d_t16.p;
    uint32_t *out = form ? d_out1.as<uint32_t>() : d_out0.as<uint32_t>();
    HIPCHK(c, msvk_forms_probe(form, d_res.as<uint8_t>(), L, tab, tab_u4, bias2, out, grid, c->stream));   // warm-up
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, msvk_forms_probe(form, d_res.as<uint8_t>(), L, tab, tab_u4, bias2, out, grid, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    ms[form] = t;
  }
  std::vector<uint32_t> o0((size_t) nw), o1((size_t) nw);
  HIPCHK(c, hipMemcpy(o0.data(), d_out0.p, (size_t) nw * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(o1.data(), d_out1.p, (size_t) nw * 4, hipMemcpyDeviceToHost));
  int32_t same = 1;
  for (int w = 0; w < nw; ++w) same &= (o0[(size_t) w] == o1[(size_t) w] && o0[(size_t) w] < 2048u) ? 1 : 0;
  *ms_f16 = ms[0];
  *ms_u16 = ms[1];
  *cells = (int64_t) nw * L * 2048;
  *agree = same;
  return MSVG_OK;
}

}  // extern "C"

namespace {

// Register class of the scan kernel for Q = ceil(M/128).
int scan_class(int Q) {
  for (int q : {2, 4, 8, 12, 16, 20, 24, 28, 32})
    if (Q <= q) return q;
  return 0;
}

// The scan set for pids: classes, per-profile tables at their class, and the
// device descriptors; reused while the same profiles are scanned.
int scan_prepare(msvg_ctx *c, const int32_t *pids, int32_t np, ScanSet **ret) {
  if (c->scanset && (int32_t) c->scanset->pids.size() == np &&
      std::equal(pids, pids + np, c->scanset->pids.begin())) {
    *ret = c->scanset.get();
    return MSVG_OK;
  }
  c->scanset.reset();
  auto set = std::make_unique<ScanSet>();
  set->pids.assign(pids, pids + np);
  set->prof.resize((size_t) np);
  for (int32_t i = 0; i < np; ++i) {
    auto it = c->profiles.find(pids[i]);
    if (it == c->profiles.end()) return fail(c, MSVG_EINVAL, "unknown profile id");
    set->prof[(size_t) i] = it->second.get();
  }
  set->Kp = set->prof[0]->Kp;
  for (int32_t i = 1; i < np; ++i)
    if (set->prof[(size_t) i]->Kp != set->Kp) return fail(c, MSVG_EINVAL, "scanned profiles must share one alphabet");
  std::map<int, ScanSet::Cls *> by_q;
  hipStream_t st = c->stream;
  if (!c->d_idrow.p) {
    uint8_t id[256];
    for (int x = 0; x < 256; ++x) id[x] = (uint8_t) x;
    HIPCHK(c, c->d_idrow.ensure(256));
    HIPCHK(c, hipMemcpy(c->d_idrow.p, id, 256, hipMemcpyHostToDevice));
  }
  for (int32_t i = 0; i < np; ++i) {
    Profile &p = *set->prof[(size_t) i];
    set->sumM += p.M;
    const int Q = scan_class(qfor(p.M, 128));
    if (Q == 0 || fast_lds(p.Kp, Q) > (size_t) kLdsBytes || p.Kp > 254) { set->fallback.push_back(i); continue; }
    if (p.scanQ != Q) {
      HIPCHK(c, p.d_scantab.ensure((size_t) (p.Kp + 1) * ((Q + 3) / 4) * 64 * 16));
      HIPCHK(c, msvk_build_fast_table(p.d_rbv.as<uint8_t>(), p.M, p.Kp, c->d_idrow.as<uint8_t>(), p.Kp, 64, Q, 64, p.bias,
                                      p.d_scantab.as<uint32_t>(), st));
      p.scanQ = Q;
    }
    auto it = by_q.find(Q);
    if (it == by_q.end()) {
      set->cls.push_back(std::make_unique<ScanSet::Cls>());
      set->cls.back()->Q = Q;
      it = by_q.emplace(Q, set->cls.back().get()).first;
    }
    it->second->members.push_back(i);
  }
  std::sort(set->cls.begin(), set->cls.end(),
            [](const std::unique_ptr<ScanSet::Cls> &x, const std::unique_ptr<ScanSet::Cls> &y) { return x->Q > y->Q; });
  for (auto &cl : set->cls) {
    std::vector<ScanProf> h(cl->members.size());
    for (size_t k = 0; k < cl->members.size(); ++k) {
      const Profile &p = *set->prof[(size_t) cl->members[k]];
      ScanProf &sp = h[k];
      std::memset(&sp, 0, sizeof sp);
      sp.tab = p.d_scantab.as<uint4>();
      sp.prof.M = p.M; sp.prof.Q = cl->Q; sp.prof.QB = (cl->Q + 3) / 4; sp.prof.Qs = qfor(p.M, 128);
      sp.prof.tbm = p.tbm; sp.prof.tec = p.tec; sp.prof.base = p.base; sp.prof.bias = p.bias;
      sp.prof.slow_all = 0;
    }
    HIPCHK(c, cl->d_profs.ensure(h.size() * sizeof(ScanProf)));
    HIPCHK(c, hipMemcpyAsync(cl->d_profs.p, h.data(), h.size() * sizeof(ScanProf), hipMemcpyHostToDevice, st));
    cl->slot.assign((size_t) np, -1);
    for (size_t k = 0; k < cl->members.size(); ++k) cl->slot[(size_t) cl->members[k]] = (int32_t) k;
    cl->h_profs = std::move(h);
  }
  HIPCHK(c, hipStreamSynchronize(st));
  c->scanset = std::move(set);
  *ret = c->scanset.get();
  return MSVG_OK;
}

// Lanes per query for the batched scan, by model length: the fast kernel's
// tuned lane counts on this device (8 up to M = 128, 16 up to 768, 32 up to
// 1024, a whole wave beyond).
int scan_lanes(int M) { return M <= 128 ? 8 : M <= 768 ? 16 : M <= 1024 ? 32 : 64; }

// LDS of a grouped-scan workgroup staging R + 1 rows of a table with QB
// register blocks: G = 64 whole rows of the fast layout; G < 64 the narrow
// layout's 32-row block stride, the last block only as far as row R.
size_t scan_group_lds(int G, int R, int QB) {
  if (G == 64) return (size_t) (R + 1) * QB * 1024;
  const size_t tw = (size_t) (G < 16 ? 16 : G);
  return ((size_t) (QB - 1) * 32 + (size_t) (R + 1)) * tw * 16;
}

// The table rows a batch of queries uses: the codes present (rowcode[r],
// r < R, ascending) and the null row R, which takes table row Kp; rowof maps
// a code to its row.  Returns R.
int scan_rows(const uint8_t *res, const int64_t *qoff, const int32_t *qlen, const std::vector<int32_t> &qs, int Kp,
              uint8_t rowof[256], std::vector<uint8_t> &rowcode) {
  bool seen[256] = {false};
  for (int32_t q : qs)
    for (int32_t k = 0; k < qlen[q]; ++k) seen[res[qoff[q] + k]] = true;
  rowcode.clear();
  for (int x = 0; x < Kp; ++x)
    if (seen[x]) { rowof[x] = (uint8_t) rowcode.size(); rowcode.push_back((uint8_t) x); }
  const int R = (int) rowcode.size();
  rowcode.push_back((uint8_t) Kp);
  return R;
}

// Grouped tables and classes of a scan set (every profile that has a
// one-query class has one: at G = 64 the grouped table is no larger).
int scan_prepare_grouped(msvg_ctx *c, ScanSet *set) {
  if (set->grouped) return MSVG_OK;
  hipStream_t st = c->stream;
  std::map<std::pair<int, int>, ScanSet::GCls *> by_tile;
  std::vector<char> has_cls(set->pids.size(), 0);
  for (auto &cl : set->cls)
    for (int32_t i : cl->members) has_cls[(size_t) i] = 1;
  for (size_t i = 0; i < set->pids.size(); ++i) {
    if (!has_cls[i]) continue;
    Profile &p = *set->prof[i];
    // the tuned lanes for M, wider where the registers or the table (over
    // the whole alphabet) would not fit; the narrow layouts take at most 32
    // rows (Kp + 1 <= 30 for every alphabet)
    int G = scan_lanes(p.M), Q = qfor(p.M, 2 * G);
    while (G < 64 && (Q > bat_qmax(G) || scan_group_lds(G, p.Kp, (Q + 3) / 4) > (size_t) kLdsBytes)) {
      G *= 2;
      Q = qfor(p.M, 2 * G);
    }
    const int QB = (Q + 3) / 4;
    if (Q > bat_qmax(G) || scan_group_lds(G, p.Kp, QB) > (size_t) kLdsBytes || !msvk_scan_group_built(G, QB) ||
        (G < 64 && p.Kp + 1 > 32))
      return fail(c, MSVG_EINVAL, "scan profile without a grouped kernel");
    if (p.gG != G || p.gQ != Q) {
      if (G == 64) {
        HIPCHK(c, p.d_gtab.ensure(fast_lds(p.Kp, Q)));
        HIPCHK(c, msvk_build_fast_table(p.d_rbv.as<uint8_t>(), p.M, p.Kp, c->d_idrow.as<uint8_t>(), p.Kp, 64, Q, 64,
                                        p.bias, p.d_gtab.as<uint32_t>(), st));
      } else {
        HIPCHK(c, p.d_gtab.ensure(bat_bytes(G, p.Kp, Q)));
        HIPCHK(c, msvk_build_bat_table(p.d_rbv.as<uint8_t>(), p.M, c->d_idrow.as<uint8_t>(), p.Kp, G, Q, p.bias,
                                       p.d_gtab.as<uint32_t>(), st));
      }
      p.gG = G; p.gQ = Q;
    }
    auto it = by_tile.find({G, QB});
    if (it == by_tile.end()) {
      set->gcls.push_back(std::make_unique<ScanSet::GCls>());
      set->gcls.back()->G = G; set->gcls.back()->QB = QB;
      it = by_tile.emplace(std::make_pair(G, QB), set->gcls.back().get()).first;
    }
    it->second->members.push_back((int32_t) i);
  }
  for (auto &gc : set->gcls) {         // most registers first: their workgroups go out first
    std::stable_sort(gc->members.begin(), gc->members.end(), [set](int32_t x, int32_t y) {
      return set->prof[(size_t) x]->gQ > set->prof[(size_t) y]->gQ;
    });
    gc->Qmax = set->prof[(size_t) gc->members[0]]->gQ;
    gc->Qsum = 0;
    for (int32_t i : gc->members) gc->Qsum += set->prof[(size_t) i]->gQ;
  }
  for (auto &gc : set->gcls) {
    std::vector<ProfDev> h(gc->members.size());
    for (size_t k = 0; k < gc->members.size(); ++k) {
      const Profile &p = *set->prof[(size_t) gc->members[k]];
      ProfDev &d = h[k];
      std::memset(&d, 0, sizeof d);
      d.ftab = p.d_gtab.as<uint4>();
      d.M = p.M; d.Q = p.gQ; d.QB = gc->QB; d.Qs = qfor(p.M, 128);
      d.tbm = p.tbm; d.tec = p.tec; d.base = p.base; d.bias = p.bias;
    }
    HIPCHK(c, gc->d_profs.ensure(h.size() * sizeof(ProfDev)));
    HIPCHK(c, hipMemcpyAsync(gc->d_profs.p, h.data(), h.size() * sizeof(ProfDev), hipMemcpyHostToDevice, st));
  }
  HIPCHK(c, hipStreamSynchronize(st));
  set->grouped = true;
  return MSVG_OK;
}

// Batches of at least this many queries take the grouped kernel.
constexpr int kScanGroupMin = 4;
constexpr int kScanCUs = 256;

// Wave-slot schedule of one grouped launch.  Task t (queries t*NG ..
// t*NG+NG-1 of the batch sorted longest first) costs its longest query's
// padded rows plus 16 (its set-up and decision); every profile of a class
// sees the same tasks, so one schedule serves all of its workgroups.
// S slots are split * waves + wave over nsplit workgroups per profile.
struct ScanPlan {
  int waves = 1, nsplit = 1;
  int64_t load = 0, sum = 0;         // rows of the busiest slot, of all tasks
  std::vector<int32_t> off, task;    // slot w's tasks: task[off[w] .. off[w + 1])
};

// Longest processing time first over S slots: each task, in order (cost is
// non-increasing), goes to the slot with the fewest rows.  Returns the
// largest slot load; fills the slot lists when plan is given.
int64_t scan_lpt(const std::vector<int64_t> &cost, int S, ScanPlan *plan) {
  using Load = std::pair<int64_t, int>;
  std::priority_queue<Load, std::vector<Load>, std::greater<Load>> h;
  for (int s = 0; s < S; ++s) h.push({0, s});
  std::vector<int32_t> slot_of(plan ? cost.size() : 0);
  int64_t mx = 0;
  for (size_t t = 0; t < cost.size(); ++t) {
    Load l = h.top();
    h.pop();
    l.first += cost[t];
    mx = std::max(mx, l.first);
    if (plan) slot_of[t] = l.second;
    h.push(l);
  }
  if (plan) {
    plan->off.assign((size_t) S + 1, 0);
    for (int32_t s : slot_of) ++plan->off[(size_t) s + 1];
    for (int s = 0; s < S; ++s) plan->off[(size_t) s + 1] += plan->off[(size_t) s];
    plan->task.assign(cost.size(), 0);
    std::vector<int32_t> fill(plan->off.begin(), plan->off.end() - 1);
    for (size_t t = 0; t < cost.size(); ++t) plan->task[(size_t) fill[(size_t) slot_of[t]]++] = (int32_t) t;
  }
  return mx;
}

// Slots per profile: the most whose schedule keeps at least 93% of the slot
// rows busy (sum of costs / (S * largest load)); a class of few profiles,
// whose waves alone would not cover the device, may go down to 85% for up
// to about four waves per CU, so its long models do not hold the batch up.
// At most 16 waves per workgroup; more slots are split over workgroups.
ScanPlan scan_plan(const std::vector<int64_t> &cost, int nprof) {
  const int nt = (int) cost.size();
  int64_t sum = 0;
  for (int64_t x : cost) sum += x;
  const int smax = std::max(1, std::min(nt, 64));
  int s_many = 1, s_few = 0;
  for (int S = smax; S >= 1; --S) {                 // largest S first; stop at the first 93% schedule
    const double eff = (double) sum / ((double) S * (double) std::max<int64_t>(1, scan_lpt(cost, S, nullptr)));
    if (s_few == 0 && eff >= 0.85) s_few = S;
    if (eff >= 0.93) { s_many = S; break; }
  }
  int S = s_many;
  if ((int64_t) nprof * s_many < 4 * kScanCUs)
    S = std::max(s_many, std::min(std::max(s_few, 1), (4 * kScanCUs + nprof - 1) / nprof));
  ScanPlan p;
  p.nsplit = (S + 15) / 16;
  p.waves = (S + p.nsplit - 1) / p.nsplit;
  p.load = scan_lpt(cost, p.waves * p.nsplit, &p);
  p.sum = sum;
  return p;
}

// Streams of a grouped scan's launches.  A stream runs its launches one
// after the other.  A class with fewer workgroups than CUs is bound by its
// busiest slot's row chain, not by the device's throughput, so such classes
// get streams of their own (two of four), longest chain (register count x
// rows) first, each to the stream whose chains add up to the least; the
// other classes share the remaining streams the same way by their work
// (profiles x registers x rows).  Returns each class's stream and fills
// `order` with the launch order and `prio` with the classes whose waves take
// a higher issue priority (the chain-bound ones, when both kinds are there:
// among the busier launches' waves their chains would stretch several-fold).
std::vector<int> scan_streams(const std::vector<std::unique_ptr<ScanSet::GCls>> &gcls,
                              const std::vector<ScanPlan> &plans, int ns, std::vector<size_t> &order,
                              std::vector<char> &prio) {
  const size_t n = gcls.size();
  std::vector<int> stream(n, 0);
  std::vector<int64_t> key(n);
  prio.assign(n, 0);
  std::vector<size_t> small, big;
  for (size_t ci = 0; ci < n; ++ci) {
    const int64_t wgs = (int64_t) gcls[ci]->members.size() * plans[ci].nsplit;
    if (wgs < kScanCUs) {
      small.push_back(ci);
      key[ci] = (int64_t) gcls[ci]->Qmax * plans[ci].load;
    } else {
      big.push_back(ci);
      key[ci] = gcls[ci]->Qsum * plans[ci].sum;
    }
  }
  auto by_key = [&key](size_t x, size_t y) { return key[x] != key[y] ? key[x] > key[y] : x < y; };
  std::sort(small.begin(), small.end(), by_key);
  std::sort(big.begin(), big.end(), by_key);
  auto deal = [&](const std::vector<size_t> &cls, int s0, int s1) {
    std::vector<int64_t> acc((size_t) (s1 - s0), 0);
    for (size_t ci : cls) {
      const int k = (int) (std::min_element(acc.begin(), acc.end()) - acc.begin());
      acc[(size_t) k] += key[ci];
      stream[ci] = s0 + k;
    }
  };
  if (small.empty() || big.empty() || ns < 4) {
    std::vector<size_t> all(small);
    all.insert(all.end(), big.begin(), big.end());
    std::sort(all.begin(), all.end(), by_key);
    deal(all, 0, ns);
    order = all;
  } else {
    deal(small, 0, ns / 2);
    deal(big, ns / 2, ns);
    for (size_t ci : small) prio[ci] = 1;
    // alternate, so each stream's first launch goes out at once, a
    // chain-bound one first: its workgroups (larger tables) then find CUs
    // with LDS free before the busier launches fill them (issued after one,
    // a four-profile class of 110 KB tables waited 1.5 ms for its CUs)
    order.clear();
    for (size_t i = 0; i < std::max(small.size(), big.size()); ++i) {
      if (i < small.size()) order.push_back(small[i]);
      if (i < big.size()) order.push_back(big[i]);
    }
  }
  return stream;
}

// msvg_scan_many for a batch: the fast kernel in scan mode (one workgroup
// per profile, 64/G queries per wave, longest queries first), then the
// J-aware kernel for the (profile, query) pairs whose J state may fire,
// listed per profile.  Pairs neither settles go to `slow`.
int scan_grouped(msvg_ctx *c, ScanSet *set, int32_t np, const uint8_t *res, const int64_t *qoff, const int32_t *qlen,
                 double F1, float *usc, uint8_t *status, uint8_t *pass, const std::vector<int32_t> &qlive,
                 std::vector<std::pair<int32_t, int32_t>> &slow) {
  int rc = scan_prepare_grouped(c, set);
  if (rc != MSVG_OK) return rc;
  const int Kp = set->Kp;
  const int32_t nl = (int32_t) qlive.size();
  std::vector<int32_t> order(qlive);
  std::stable_sort(order.begin(), order.end(), [qlen](int32_t x, int32_t y) { return qlen[x] > qlen[y]; });
  int64_t padded = 0;
  for (int32_t q : order) padded += ((int64_t) qlen[q] + 15) & ~(int64_t) 15;
  uint8_t rowof[256];
  std::vector<uint8_t> rowcode;
  const int R = scan_rows(res, qoff, qlen, order, Kp, rowof, rowcode);
  // staging: row indices of the codes (only the rows the batch uses are
  // staged in LDS) padded to 16 with the null row R, then offsets, lengths,
  // tjb_b and the row list
  const size_t o_off = ((size_t) padded + 15) & ~(size_t) 15;
  const size_t o_len = o_off + (size_t) nl * sizeof(int64_t);
  const size_t o_tjb = o_len + (size_t) nl * sizeof(int32_t);
  const size_t o_rc = o_tjb + (size_t) nl;
  // the wave-slot schedules, one per class, after the row list
  std::vector<ScanPlan> plans;
  std::vector<size_t> o_plan;
  size_t o_sched = (o_rc + rowcode.size() + 15) & ~(size_t) 15;
  {
    std::map<int, std::vector<int64_t>> costs;                // per G
    for (auto &gc : set->gcls) {
      std::vector<int64_t> &cost = costs[gc->G];
      if (cost.empty()) {
        const int NG = 64 / gc->G;
        for (int32_t k = 0; k < nl; k += NG) cost.push_back((((int64_t) qlen[order[(size_t) k]] + 15) & ~(int64_t) 15) + 16);
      }
      plans.push_back(scan_plan(cost, (int) gc->members.size()));
      o_plan.push_back(o_sched);
      o_sched += ((plans.back().off.size() + plans.back().task.size()) * sizeof(int32_t) + 15) & ~(size_t) 15;
    }
  }
  const size_t total = o_sched;
  if (c->h_one_cap < total + 64) {
    if (c->h_one_res) { (void) hipHostFree(c->h_one_res); c->h_one_res = nullptr; c->h_one_cap = 0; }
    const size_t cap = std::max<size_t>(total + 64, 4096);
    HIPCHK(c, hipHostMalloc((void **) &c->h_one_res, cap, hipHostMallocDefault));
    c->h_one_cap = cap;
  }
  uint8_t *h = c->h_one_res;
  int64_t *h_off = reinterpret_cast<int64_t *>(h + o_off);
  int32_t *h_len = reinterpret_cast<int32_t *>(h + o_len);
  uint8_t *h_tjb = h + o_tjb;
  std::memcpy(h + o_rc, rowcode.data(), rowcode.size());
  for (size_t ci = 0; ci < plans.size(); ++ci) {
    const ScanPlan &pl = plans[ci];
    std::memcpy(h + o_plan[ci], pl.off.data(), pl.off.size() * sizeof(int32_t));
    std::memcpy(h + o_plan[ci] + pl.off.size() * sizeof(int32_t), pl.task.data(), pl.task.size() * sizeof(int32_t));
  }
  std::vector<int> tjb_k((size_t) nl);
  int64_t pos = 0;
  for (int32_t k = 0; k < nl; ++k) {
    const int32_t q = order[(size_t) k];
    const int64_t Lq = qlen[q], Lpad = (Lq + 15) & ~(int64_t) 15;
    const uint8_t *src = res + qoff[q];
    for (int64_t i = 0; i < Lq; ++i) h[pos + i] = rowof[src[i]];
    std::memset(h + pos + Lq, R, (size_t) (Lpad - Lq));
    h_off[k] = pos; h_len[k] = (int32_t) Lq;
    tjb_k[(size_t) k] = msvh_tjb((int) Lq);
    h_tjb[k] = (uint8_t) tjb_k[(size_t) k];
    pos += Lpad;
  }
  size_t nmem = 0;
  for (auto &gc : set->gcls) nmem += gc->members.size();
  const size_t nw = nmem * (size_t) nl;                      // uint16 result words
  hipStream_t st = c->stream;
  HIPCHK(c, c->d_one_res.ensure(std::max<size_t>(total, 4096)));
  HIPCHK(c, c->d_scanout.ensure(std::max<size_t>(nw * sizeof(uint16_t), 16)));
  if (c->h_scancap * sizeof(uint32_t) < nw * sizeof(uint16_t)) {
    if (c->h_scanout) { (void) hipHostFree(c->h_scanout); c->h_scanout = nullptr; c->h_scancap = 0; }
    const size_t cap = nw / 2 + 1;
    HIPCHK(c, hipHostMalloc((void **) &c->h_scanout, cap * sizeof(uint32_t), hipHostMallocDefault));
    c->h_scancap = cap;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_one_res.p, h, total, hipMemcpyHostToDevice, st));
  const uint8_t *d = c->d_one_res.as<uint8_t>();
  DbDev qdb{};
  qdb.res = d;
  qdb.off = reinterpret_cast<const int64_t *>(d + o_off);
  qdb.len = reinterpret_cast<const int32_t *>(d + o_len);
  qdb.tjb = d + o_tjb;
  qdb.n = nl;
  qdb.R = R;
  uint16_t *d_words = c->d_scanout.as<uint16_t>();
  HIPCHK(c, hipEventRecord(c->ev[0], st));
  const int ns = std::min<int>(kVitStreams, (int) set->gcls.size());
  if (ns > 1) {
    if (!c->vstream[0]) {
      for (int k = 0; k < kVitStreams; ++k) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->vstream[k], hipStreamNonBlocking));
        HIPCHK(c, hipEventCreateWithFlags(&c->vdone[k], hipEventDisableTiming));
      }
    }
    HIPCHK(c, hipEventRecord(c->fork, st));
    for (int k = 0; k < ns; ++k) HIPCHK(c, hipStreamWaitEvent(c->vstream[k], c->fork, 0));
  }
  std::vector<size_t> cbase(set->gcls.size(), 0), lorder;       // each class's first result row
  for (size_t ci = 1; ci < set->gcls.size(); ++ci) cbase[ci] = cbase[ci - 1] + set->gcls[ci - 1]->members.size();
  std::vector<char> cprio;
  const std::vector<int> cstream = scan_streams(set->gcls, plans, ns, lorder, cprio);
  for (size_t ci : lorder) {
    auto &gc = set->gcls[ci];
    const ScanPlan &pl = plans[ci];
    const int nprof = (int) gc->members.size();
    ScanGroupArgs a{};
    a.profs = gc->d_profs.as<ProfDev>();
    a.np = nprof;
    a.db = qdb;
    a.rowcode = d + o_rc;
    a.soff = reinterpret_cast<const int32_t *>(d + o_plan[ci]);
    a.stask = a.soff + pl.off.size();
    a.out = d_words + cbase[ci];                            // words [sorted query][row]
    a.stride = (int64_t) nmem;
    a.prio = cprio[ci];
    hipStream_t cs = ns > 1 ? c->vstream[cstream[ci]] : st;
    HIPCHK(c, msvk_launch_scan_group(gc->G, gc->QB, &a, nprof * pl.nsplit, 64 * pl.waves,
                                     scan_group_lds(gc->G, R, gc->QB), cs));
  }
  if (ns > 1) {
    for (int k = 0; k < ns; ++k) {
      HIPCHK(c, hipEventRecord(c->vdone[k], c->vstream[k]));
      HIPCHK(c, hipStreamWaitEvent(st, c->vdone[k], 0));
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[1], st));
  HIPCHK(c, hipMemcpyAsync(c->h_scanout, d_words, nw * sizeof(uint16_t), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->stats.ms_fast = ms;

  std::vector<PassRule> rules;
  rules.reserve((size_t) nl);
  for (int32_t k = 0; k < nl; ++k) rules.emplace_back(h_len[k], F1);
  const uint16_t *words = reinterpret_cast<const uint16_t *>(c->h_scanout);
  auto settle = [&](int32_t i, int32_t k, uint32_t w) {   // a final word of profile position i, sorted query k
    const int32_t q = order[(size_t) k];
    const size_t o = (size_t) q * np + i;
    const Profile &p = *set->prof[(size_t) i];
    if ((w >> 8) == RES_ERANGE) {
      usc[o] = std::numeric_limits<float>::infinity(); status[o] = MSVG_ERANGE; pass[o] = 1;
    } else {
      usc[o] = usc_of((int) (w & 255u) - tjb_k[(size_t) k], p.base, p.scale);
      status[o] = MSVG_OK;
      pass[o] = rules[(size_t) k](usc[o], p.mu, p.lambda) ? 1 : 0;
    }
  };
  std::vector<std::vector<std::pair<int32_t, int32_t>>> jpairs(set->cls.size());   // per class: (slot, k)
  std::vector<int> cls_of(set->pids.size(), -1);
  for (size_t ci = 0; ci < set->cls.size(); ++ci)
    for (int32_t i : set->cls[ci]->members) cls_of[(size_t) i] = (int) ci;
  // The words are decoded on host threads, each over a range of queries:
  // word (k, r) at words[k * nmem + r], row r = a class member, and its
  // outputs at [query][profile position], so reads and writes run along rows.
  if (!set->grows_ok) {
    set->grows.clear();
    for (auto &gc : set->gcls)
      for (int32_t i : gc->members) {
        const Profile &p = *set->prof[(size_t) i];
        set->grows.push_back(ScanSet::Row{i, p.base, p.scale, p.mu, p.lambda});
      }
    set->grows_ok = true;
  }
  const std::vector<ScanSet::Row> &rows = set->grows;
  const int nth = (int) std::max<size_t>(1, std::min<size_t>({(size_t) 8, (size_t) nl,
                                                             nmem * (size_t) nl / 32768 + 1,
                                                             (size_t) std::max(1u, std::thread::hardware_concurrency())}));
  struct Local { std::vector<std::array<int32_t, 3>> j; std::vector<std::pair<int32_t, int32_t>> s; };
  std::vector<Local> loc((size_t) nth);
  auto work = [&](int t) {
    const int32_t k0 = (int32_t) ((int64_t) nl * t / nth), k1 = (int32_t) ((int64_t) nl * (t + 1) / nth);
    Local &L = loc[(size_t) t];
    for (int32_t k = k0; k < k1; ++k) {
      const int32_t q = order[(size_t) k];
      const int tjb = tjb_k[(size_t) k];
      const PassRule &rule = rules[(size_t) k];
      const uint16_t *wk = words + (size_t) k * nmem;
      float *uq = usc + (size_t) q * np;
      uint8_t *sq = status + (size_t) q * np, *pq = pass + (size_t) q * np;
      for (size_t r = 0; r < nmem; ++r) {
        const uint32_t w = wk[r];
        const ScanSet::Row &R = rows[r];
        const uint32_t code = w >> 8;
        if (code == RES_SLOW) {
          if (w & 1u) {
            const int ci = cls_of[(size_t) R.i];
            L.j.push_back({ci, set->cls[(size_t) ci]->slot[(size_t) R.i], k});
          } else {
            L.s.emplace_back(R.i, q);
          }
        } else if (code == RES_ERANGE) {
          uq[R.i] = std::numeric_limits<float>::infinity(); sq[R.i] = MSVG_ERANGE; pq[R.i] = 1;
        } else {
          const float u = usc_of((int) (w & 255u) - tjb, R.base, R.scale);
          uq[R.i] = u; sq[R.i] = MSVG_OK; pq[R.i] = rule(u, R.mu, R.lambda) ? 1 : 0;
        }
      }
    }
  };
  {
    std::vector<std::thread> th;
    int t = 1;
    try {
      for (; t < nth; ++t) th.emplace_back(work, t);
    } catch (...) {                    // a thread that cannot start: its queries are decoded here
    }
    for (int u = t; u < nth; ++u) work(u);
    work(0);
    for (auto &x : th) x.join();
  }
  size_t nj = 0;
  for (const Local &L : loc) {
    for (const auto &e : L.j) jpairs[(size_t) e[0]].emplace_back(e[1], e[2]);
    nj += L.j.size();
    slow.insert(slow.end(), L.s.begin(), L.s.end());
  }
  c->stats.n_jrow = (int64_t) nj;
  if (nj == 0) return MSVG_OK;

  // J stage: per one-query class, the profiles with such pairs, each
  // workgroup over its listed queries.
  struct Part { size_t ci, o_prof, o_woff, o_widx, o_out, nwg, nit; };
  std::vector<Part> parts;
  std::vector<uint8_t> hb;
  std::vector<std::pair<int32_t, int32_t>> items;            // (profile position, k) per output word
  size_t nout = 0;
  for (size_t ci = 0; ci < set->cls.size(); ++ci) {
    auto &jp = jpairs[ci];
    if (jp.empty()) continue;
    std::sort(jp.begin(), jp.end());
    const auto &cl = *set->cls[ci];
    std::vector<ScanProf> profs;
    std::vector<int32_t> woff{0}, widx;
    for (size_t t = 0; t < jp.size(); ++t) {
      if (t == 0 || jp[t].first != jp[t - 1].first) {
        if (t > 0) woff.push_back((int32_t) widx.size());
        profs.push_back(cl.h_profs[(size_t) jp[t].first]);
      }
      widx.push_back(jp[t].second);
      items.emplace_back(cl.members[(size_t) jp[t].first], jp[t].second);
    }
    woff.push_back((int32_t) widx.size());
    Part pt{};
    pt.ci = ci; pt.nwg = profs.size(); pt.nit = widx.size();
    auto put = [&hb](const void *src, size_t n) {
      const size_t o = (hb.size() + 15) & ~(size_t) 15;
      hb.resize(o + n);
      std::memcpy(hb.data() + o, src, n);
      return o;
    };
    pt.o_prof = put(profs.data(), profs.size() * sizeof(ScanProf));
    pt.o_woff = put(woff.data(), woff.size() * sizeof(int32_t));
    pt.o_widx = put(widx.data(), widx.size() * sizeof(int32_t));
    pt.o_out = nout;
    nout += widx.size();
    parts.push_back(pt);
  }
  const size_t o_jout = (hb.size() + 15) & ~(size_t) 15;
  HIPCHK(c, c->d_scanj.ensure(o_jout + nout * sizeof(uint32_t)));
  HIPCHK(c, hipMemcpyAsync(c->d_scanj.p, hb.data(), hb.size(), hipMemcpyHostToDevice, st));
  uint8_t *dj = c->d_scanj.as<uint8_t>();
  HIPCHK(c, hipEventRecord(c->ev[2], st));
  // the classes' J-aware launches side by side on the scan's streams
  const int nj_st = std::min<int>(kVitStreams, (int) parts.size());
  if (nj_st > 1) {
    if (!c->vstream[0]) {
      for (int k = 0; k < kVitStreams; ++k) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->vstream[k], hipStreamNonBlocking));
        HIPCHK(c, hipEventCreateWithFlags(&c->vdone[k], hipEventDisableTiming));
      }
    }
    HIPCHK(c, hipEventRecord(c->fork, st));
    for (int k = 0; k < nj_st; ++k) HIPCHK(c, hipStreamWaitEvent(c->vstream[k], c->fork, 0));
  }
  size_t jpart = 0;
  for (const Part &pt : parts) {
    ScanArgs a{};
    a.profs = reinterpret_cast<const ScanProf *>(dj + pt.o_prof);
    a.res = d;
    a.qoff = qdb.off; a.qlen = qdb.len; a.qtjb = qdb.tjb;
    a.nq = nl; a.R = R; a.stride = 0;
    a.rowcode = d + o_rc;
    a.out = reinterpret_cast<uint32_t *>(dj + o_jout) + pt.o_out;
    a.wq_off = reinterpret_cast<const int32_t *>(dj + pt.o_woff);
    a.wq_idx = reinterpret_cast<const int32_t *>(dj + pt.o_widx);
    const int Q = set->cls[pt.ci]->Q;
    hipStream_t js = nj_st > 1 ? c->vstream[jpart++ % (size_t) nj_st] : st;
    HIPCHK(c, msvk_launch_scan(Q, &a, (int) pt.nwg, fast_lds(R, Q), js));
  }
  if (nj_st > 1) {
    for (int k = 0; k < nj_st; ++k) {
      HIPCHK(c, hipEventRecord(c->vdone[k], c->vstream[k]));
      HIPCHK(c, hipStreamWaitEvent(st, c->vdone[k], 0));
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[3], st));
  std::vector<uint32_t> jw(nout);
  HIPCHK(c, hipMemcpyAsync(jw.data(), dj + o_jout, nout * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
  c->stats.ms_fast += ms;
  for (size_t t = 0; t < nout; ++t) {
    const uint32_t w = jw[t];
    if ((w >> 8) == RES_SLOW) slow.emplace_back(items[t].first, order[(size_t) items[t].second]);
    else settle(items[t].first, items[t].second, w);
  }
  return MSVG_OK;
}

}  // namespace

extern "C" {

int msvg_scan(msvg_ctx *c, int32_t np, const int32_t *pids, const uint8_t *dsq, int L, double F1, float *usc,
              uint8_t *status, uint8_t *pass) {
  if (!c) return MSVG_EINVAL;
  if (L < 0 || (np > 0 && !dsq)) return fail(c, MSVG_EINVAL, "null argument");
  const int64_t off = 0;
  const int32_t len = L;
  static const uint8_t none = 0;
  return msvg_scan_many(c, np, pids, 1, L > 0 ? dsq + 1 : &none, &off, &len, F1, usc, status, pass);
}

int msvg_scan_many(msvg_ctx *c, int32_t np, const int32_t *pids, int32_t nq, const uint8_t *res, const int64_t *qoff,
                   const int32_t *qlen, double F1, float *usc, uint8_t *status, uint8_t *pass) {
  if (!c) return MSVG_EINVAL;
  if (np < 0 || nq < 0 || (np > 0 && nq > 0 && (!pids || !res || !qoff || !qlen || !usc || !status || !pass)))
    return fail(c, MSVG_EINVAL, "null argument");
  c->stats = msvg_stats{};
  if (np == 0 || nq == 0) return MSVG_OK;
  for (int32_t q = 0; q < nq; ++q)
    if (qlen[q] < 0 || qoff[q] < 0) return fail(c, MSVG_EINVAL, "negative query offset or length");
  (void) hipSetDevice(c->device);
  ScanSet *set;
  int rc = scan_prepare(c, pids, np, &set);        // validates ids and alphabet when it builds
  if (rc != MSVG_OK) return rc;
  const int Kp = set->Kp;
  int64_t padded = 0, residues = 0;
  for (int32_t q = 0; q < nq; ++q) {
    for (int32_t i = 0; i < qlen[q]; ++i)
      if (res[qoff[q] + i] >= Kp) return fail(c, MSVG_EINVAL, "query holds residue codes outside the profiles' alphabet");
    padded += ((int64_t) qlen[q] + 15) & ~(int64_t) 15;
    residues += qlen[q];
  }
  hipStream_t st = c->stream;
  // (profile position, query) pairs the kernel leaves: all of an empty
  // query, the profiles whose table does not fit LDS, and RES_SLOW words
  std::vector<std::pair<int32_t, int32_t>> slow;
  std::vector<int32_t> qlive;                       // queries with residues
  for (int32_t q = 0; q < nq; ++q) {
    if (qlen[q] == 0) {
      for (int32_t i = 0; i < np; ++i) slow.emplace_back(i, q);
    } else {
      qlive.push_back(q);
      for (int32_t i : set->fallback) slow.emplace_back(i, q);
    }
  }
  const int32_t nl = (int32_t) qlive.size();
  if (nl >= kScanGroupMin && !set->cls.empty()) {
    if ((rc = scan_grouped(c, set, np, res, qoff, qlen, F1, usc, status, pass, qlive, slow)) != MSVG_OK) return rc;
    c->stats.ms_total = c->stats.ms_fast;
  } else if (nl > 0) {
    // stage the live queries: row indices of their codes (only the rows
    // they use are staged in LDS) padded to 16 bytes with the null row R,
    // then offsets, lengths, tjb_b and the row list, in one pinned buffer
    uint8_t rowof[256];
    std::vector<uint8_t> rowcode;
    const int R = scan_rows(res, qoff, qlen, qlive, Kp, rowof, rowcode);
    const size_t meta = (size_t) nl * (sizeof(int64_t) + sizeof(int32_t) + 1) + rowcode.size();
    const size_t need = (size_t) padded + meta + 64;
    if (c->h_one_cap < need) {
      if (c->h_one_res) { (void) hipHostFree(c->h_one_res); c->h_one_res = nullptr; c->h_one_cap = 0; }
      const size_t cap = std::max<size_t>(need, 4096);
      HIPCHK(c, hipHostMalloc((void **) &c->h_one_res, cap, hipHostMallocDefault));
      c->h_one_cap = cap;
    }
    const size_t nout = (size_t) np * nl;
    if (c->h_scancap < nout) {
      if (c->h_scanout) { (void) hipHostFree(c->h_scanout); c->h_scanout = nullptr; c->h_scancap = 0; }
      HIPCHK(c, hipHostMalloc((void **) &c->h_scanout, nout * sizeof(uint32_t), hipHostMallocDefault));
      c->h_scancap = nout;
    }
    uint8_t *h = c->h_one_res;
    const size_t o_off = ((size_t) padded + 7) & ~(size_t) 7;
    int64_t *h_off = reinterpret_cast<int64_t *>(h + o_off);
    int32_t *h_len = reinterpret_cast<int32_t *>(h + o_off + (size_t) nl * sizeof(int64_t));
    uint8_t *h_tjb = h + o_off + (size_t) nl * (sizeof(int64_t) + sizeof(int32_t));
    uint8_t *h_rc = h_tjb + nl;
    std::memcpy(h_rc, rowcode.data(), rowcode.size());
    int64_t pos = 0;
    for (int32_t k = 0; k < nl; ++k) {
      const int32_t q = qlive[(size_t) k];
      const int64_t Lq = qlen[q], Lpad = (Lq + 15) & ~(int64_t) 15;
      const uint8_t *src = res + qoff[q];
      for (int64_t i = 0; i < Lq; ++i) h[pos + i] = rowof[src[i]];
      std::memset(h + pos + Lq, R, (size_t) (Lpad - Lq));
      h_off[k] = pos;
      h_len[k] = (int32_t) Lq;
      h_tjb[k] = (uint8_t) msvh_tjb((int) Lq);
      pos += Lpad;
    }
    const size_t total = o_off + meta;
    HIPCHK(c, c->d_one_res.ensure(std::max<size_t>(total, 4096)));
    HIPCHK(c, c->d_scanout.ensure(nout * sizeof(uint32_t)));
    HIPCHK(c, hipMemcpyAsync(c->d_one_res.p, h, total, hipMemcpyHostToDevice, st));
    uint8_t *d = c->d_one_res.as<uint8_t>();
    HIPCHK(c, hipEventRecord(c->ev[0], st));
    // The classes' launches are independent, and each lasts at least as long
    // as one profile's row chain over a query (longer for larger Q): they
    // go round-robin over the context's extra streams, largest Q first.
    const int ns = std::min<int>(kVitStreams, (int) set->cls.size());
    if (ns > 1) {
      if (!c->vstream[0]) {
        for (int k = 0; k < kVitStreams; ++k) {
          HIPCHK(c, hipStreamCreateWithFlags(&c->vstream[k], hipStreamNonBlocking));
          HIPCHK(c, hipEventCreateWithFlags(&c->vdone[k], hipEventDisableTiming));
        }
      }
      HIPCHK(c, hipEventRecord(c->fork, st));
      for (int k = 0; k < ns; ++k) HIPCHK(c, hipStreamWaitEvent(c->vstream[k], c->fork, 0));
    }
    int64_t base = 0;
    for (size_t ci = 0; ci < set->cls.size(); ++ci) {
      auto &cl = set->cls[ci];
      ScanArgs a{};
      a.profs = cl->d_profs.as<ScanProf>();
      a.res = d;
      a.qoff = reinterpret_cast<const int64_t *>(d + o_off);
      a.qlen = reinterpret_cast<const int32_t *>(d + o_off + (size_t) nl * sizeof(int64_t));
      a.qtjb = d + o_off + (size_t) nl * (sizeof(int64_t) + sizeof(int32_t));
      a.nq = nl; a.R = R;
      a.rowcode = d + (h_rc - h);
      a.stride = np;
      a.out = c->d_scanout.as<uint32_t>() + base;
      hipStream_t cs = ns > 1 ? c->vstream[ci % ns] : st;
      HIPCHK(c, msvk_launch_scan(cl->Q, &a, (int) cl->members.size(), fast_lds(R, cl->Q), cs));
      base += (int64_t) cl->members.size();
    }
    if (ns > 1) {
      for (int k = 0; k < ns; ++k) {
        HIPCHK(c, hipEventRecord(c->vdone[k], c->vstream[k]));
        HIPCHK(c, hipStreamWaitEvent(st, c->vdone[k], 0));
      }
    }
    HIPCHK(c, hipEventRecord(c->ev[1], st));
    HIPCHK(c, hipMemcpyAsync(c->h_scanout, c->d_scanout.p, nout * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.ms_fast = ms;
    c->stats.ms_total = ms;
    if (!set->rows_ok) {
      set->rows.clear();
      for (auto &cl : set->cls)
        for (int32_t i : cl->members) {
          const Profile &p = *set->prof[(size_t) i];
          set->rows.push_back(ScanSet::Row{i, p.base, p.scale, p.mu, p.lambda});
        }
      set->rows_ok = true;
    }
    for (int32_t k = 0; k < nl; ++k) {
      const int32_t q = qlive[(size_t) k];
      const int tjb = msvh_tjb(qlen[q]);
      const PassRule pass_rule(qlen[q], F1);
      const uint32_t *wq = c->h_scanout + (size_t) k * np;
      float *uq = usc + (size_t) q * np;
      uint8_t *sq = status + (size_t) q * np, *pq = pass + (size_t) q * np;
      for (size_t m = 0; m < set->rows.size(); ++m) {
        const ScanSet::Row &R = set->rows[m];
        const uint32_t w = wq[m];
        const uint32_t code = w >> 8;
        if (code == RES_SLOW) { slow.emplace_back(R.i, q); continue; }
        if (code == RES_ERANGE) {
          uq[R.i] = std::numeric_limits<float>::infinity(); sq[R.i] = MSVG_ERANGE; pq[R.i] = 1;
        } else {
          uq[R.i] = usc_of((int) (w & 255u) - tjb, R.base, R.scale);
          sq[R.i] = MSVG_OK;
          pq[R.i] = pass_rule(uq[R.i], R.mu, R.lambda) ? 1 : 0;
        }
      }
    }
  }
  c->stats.cells = set->sumM * residues;
  c->stats.n_slow = (int64_t) slow.size();
  if (!slow.empty()) {
    // the search path, query by query on a one-target database, for the
    // pairs the kernel leaves
    std::sort(slow.begin(), slow.end(), [](const std::pair<int32_t, int32_t> &x, const std::pair<int32_t, int32_t> &y) {
      return x.second != y.second ? x.second < y.second : x.first < y.first;
    });
    const msvg_stats keep = c->stats;
    for (size_t a0 = 0; a0 < slow.size();) {
      const int32_t q = slow[a0].second;
      size_t a1 = a0;
      while (a1 < slow.size() && slow[a1].second == q) ++a1;
      if (c->scratch_db >= 0) msvg_db_remove(c, c->scratch_db);
      c->scratch_db = -1;
      const int64_t zero = 0;
      const int32_t Lq = qlen[q];
      static const uint8_t none = 0;
      int32_t id;
      if ((rc = msvg_db_add(c, Lq > 0 ? res + qoff[q] : &none, &zero, &Lq, 1, &id)) != MSVG_OK) return rc;
      c->scratch_db = id;
      std::vector<int32_t> ids(a1 - a0);
      for (size_t k = a0; k < a1; ++k) ids[k - a0] = pids[slow[k].first];
      std::vector<float> u(ids.size());
      std::vector<uint8_t> s2(ids.size()), p2(ids.size());
      if ((rc = msvg_search(c, ids.data(), (int32_t) ids.size(), id, F1, MSVG_MODE_PLAIN, u.data(), s2.data(),
                            p2.data())) != MSVG_OK)
        return rc;
      for (size_t k = a0; k < a1; ++k) {
        const size_t o = (size_t) q * np + slow[k].first;
        usc[o] = u[k - a0]; status[o] = s2[k - a0]; pass[o] = p2[k - a0];
      }
      a0 = a1;
    }
    c->stats = keep;
  }
  return MSVG_OK;
}

}  // extern "C"

namespace {

// The per-call drop-in's own path: one launch of msv_one over the target,
// staged through pinned memory, with a table over the profile's whole
// alphabet built once per profile.  *done is false (and nothing is
// reported) when the pair needs what only the batch path has: an empty
// target, codes outside the alphabet (the batch path reports them), a
// table that does not fit LDS, or the guard regime's exact sub-B0 maximum.
int one_target(msvg_ctx *c, Profile &p, const uint8_t *dsq, int L, float *ret_sc, bool *done) {
  *done = false;
  const int Q = qfor(p.M, 128), QB = (Q + 3) / 4;
  const size_t lds = fast_lds(p.Kp, Q);
  if (L == 0 || Q > kQMax || lds > (size_t) kLdsBytes || p.Kp > 254) return MSVG_OK;
  for (int i = 1; i <= L; ++i)
    if (dsq[i] >= p.Kp) return MSVG_OK;
  hipStream_t st = c->stream;
  if (!p.d_onetab.p) {
    if (!c->d_idrow.p) {
      uint8_t id[256];
      for (int x = 0; x < 256; ++x) id[x] = (uint8_t) x;
      HIPCHK(c, c->d_idrow.ensure(256));
      HIPCHK(c, hipMemcpyAsync(c->d_idrow.p, id, 256, hipMemcpyHostToDevice, st));
      HIPCHK(c, hipStreamSynchronize(st));
    }
    HIPCHK(c, p.d_onetab.ensure((size_t) (p.Kp + 1) * QB * 64 * 16));
    HIPCHK(c, msvk_build_fast_table(p.d_rbv.as<uint8_t>(), p.M, p.Kp, c->d_idrow.as<uint8_t>(), p.Kp, 64, Q, 64, p.bias,
                                    p.d_onetab.as<uint32_t>(), st));
  }
  const size_t Lpad = ((size_t) L + 15) & ~(size_t) 15;
  if (c->h_one_cap < Lpad) {
    if (c->h_one_res) { (void) hipHostFree(c->h_one_res); c->h_one_res = nullptr; c->h_one_cap = 0; }
    const size_t cap = std::max<size_t>(Lpad, 4096);
    HIPCHK(c, hipHostMalloc((void **) &c->h_one_res, cap, hipHostMallocDefault));
    c->h_one_cap = cap;
  }
  if (!c->h_one_out) HIPCHK(c, hipHostMalloc((void **) &c->h_one_out, sizeof(uint32_t), hipHostMallocDefault));
  HIPCHK(c, c->d_one_res.ensure(std::max<size_t>(Lpad, 4096)));
  HIPCHK(c, c->d_one_out.ensure(sizeof(uint32_t)));
  std::memcpy(c->h_one_res, dsq + 1, (size_t) L);
  std::memset(c->h_one_res + L, p.Kp, Lpad - (size_t) L);
  HIPCHK(c, hipMemcpyAsync(c->d_one_res.p, c->h_one_res, Lpad, hipMemcpyHostToDevice, st));
  OneArgs a{};
  a.prof.M = p.M; a.prof.Q = Q; a.prof.QB = QB; a.prof.Qs = qfor(p.M, 128);
  a.prof.tbm = p.tbm; a.prof.tec = p.tec; a.prof.base = p.base; a.prof.bias = p.bias;
  a.prof.slow_all = 0;
  a.tab = p.d_onetab.as<uint4>();
  a.res = c->d_one_res.as<uint8_t>();
  a.L = L;
  a.tjb = msvh_tjb(L);
  a.R = p.Kp;
  a.out = c->d_one_out.as<uint32_t>();
  HIPCHK(c, msvk_launch_one(Q, &a, lds, st));
  HIPCHK(c, hipMemcpyAsync(c->h_one_out, c->d_one_out.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  const uint32_t w = *c->h_one_out;
  const uint32_t code = w >> 8;
  if (code == RES_SLOW) return MSVG_OK;
  *done = true;
  c->stats = msvg_stats{};
  c->stats.cells = (int64_t) p.M * L;
  c->stats.dropin_path = 1;
  if (code == RES_ERANGE) { *ret_sc = std::numeric_limits<float>::infinity(); return MSVG_ERANGE; }
  *ret_sc = usc_of((int) (w & 255u) - a.tjb, p.base, p.scale);
  return MSVG_OK;
}

}  // namespace

extern "C" {

int msvg_MSVFilter(msvg_ctx *c, int32_t pid, const uint8_t *dsq, int L, float *ret_sc) {
  if (!c || !dsq || !ret_sc || L < 0) return fail(c, MSVG_EINVAL, "null argument");
  if (!c->profiles.count(pid)) return fail(c, MSVG_EINVAL, "unknown profile id");
  bool done = false;
  const int rc1 = one_target(c, *c->profiles[pid], dsq, L, ret_sc, &done);
  if (done || rc1 != MSVG_OK) return rc1;
  // One-target database holding dsq[1..L]; reused by the next call.
  if (c->scratch_db >= 0) msvg_db_remove(c, c->scratch_db);
  c->scratch_db = -1;
  const int64_t off = 0;
  int32_t id;
  int rc = msvg_db_add(c, dsq + 1, &off, &L, 1, &id);
  if (rc != MSVG_OK) return rc;
  c->scratch_db = id;
  float usc;
  uint8_t st;
  rc = msvg_search(c, &pid, 1, id, 0.02, MSVG_MODE_PLAIN, &usc, &st, nullptr);
  if (rc != MSVG_OK) return rc;
  c->stats.dropin_path = 2;
  *ret_sc = usc;
  return st == 16 ? MSVG_ERANGE : MSVG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving: continuous batching over one resident database.
// ---------------------------------------------------------------------------
extern "C" hipError_t msvk_compact(const uint8_t *pass, const float *usc, int64_t n, int32_t *count, int32_t *idx,
                                   float *out_usc, int32_t cap, hipStream_t st);

namespace {

struct ServerQuery {
  int64_t ticket = 0;
  std::string key;                // profile-cache key (set when admitted)
  msvg_oprofile desc{};
  std::vector<uint8_t> rbv;
  double t_submit = 0, t_admit = -1, t_done = -1;
  int state = 0;                  // 0 queued, 1 on the GPU, 2 done
  int status = MSVG_OK;
  int slot = -1;
  int64_t cells = 0;
  int batch = 0;                  // queries in the search that served this one
  size_t lds = 0, dev = 0;
  std::vector<int32_t> idx;
  std::vector<float> usc;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  Bufs bufs;
  DevMem d_cidx, d_cusc, d_ccount;          // compacted pass lists, one row of n per query of the group
  int32_t *h_count = nullptr;               // pinned: passing targets per query of the group
  std::vector<ServerQuery *> group;         // the queries of the search on this slot (empty: free)
};

}  // namespace

struct msvg_server {
  msvg_ctx *ctx = nullptr;
  int db = -1;
  double F1 = 0.02;
  int64_t hbm_budget = 0;
  int sched = 0;                  // 0 persistent launches, 2 one chunk per workgroup
  int batch_max = 1;              // queries folded into one search (MSVG_SERVE_BATCHED: kServerBatch)
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<ServerQuery *> pending;
  std::map<int64_t, std::unique_ptr<ServerQuery>> queries;
  // Query profiles added to the context, by content: reused by later
  // queries of the same profile; at most kServerProfiles are kept, the least
  // recently used one that no running query holds is removed first.
  struct CachedProfile { int pid = -1; int refs = 0; uint64_t used = 0; };
  std::map<std::string, CachedProfile> prof_cache;
  uint64_t use_clock = 0;
  int64_t cached = 0;             // prof_cache.size(), readable under mu
  std::thread th;
  std::atomic<bool> stopping{false};
  std::chrono::steady_clock::time_point t0;
  int64_t next_ticket = 1;
  msvg_server_info st{};
  double busy_since = -1;
  size_t lds_in_use = 0, dev_in_use = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

constexpr size_t kServerProfiles = 256;   // query profiles a server keeps on the device
constexpr int kServerBatch = 16;          // queries one batched-serving search takes at most

std::string profile_key(const ServerQuery &q) {
  std::string k(reinterpret_cast<const char *>(q.rbv.data()), q.rbv.size());
  char buf[128];
  std::snprintf(buf, sizeof buf, "|%d|%d|%d|%d|%d|%d|%a|%a|%a", q.desc.M, q.desc.Kp, q.desc.tbm_b, q.desc.tec_b,
                q.desc.base_b, q.desc.bias_b, q.desc.scale_b, q.desc.mmu, q.desc.mlambda);
  return k + buf;
}

// LDS a query's fast-stage workgroups hold while it runs.
size_t query_lds(msvg_ctx *c, const Db &d, int M) {
  int G = 64, w = 0;
  choose_tile(c, M, d, G, w);
  const int Q = qfor(M, 2 * G);
  if (Q > kQMax || fast_lds(d.R, Q) > (size_t) kLdsBytes) return fast_lds(d.R, qfor(std::min(M, 128 * seg_q(d.R)), 128));
  return fast_lds(d.R, Q);
}

// The query's profile leaves the cache's hold list (its search is over or
// never started).
void server_release(msvg_server *s, ServerQuery &q) {
  auto it = s->prof_cache.find(q.key);
  if (it != s->prof_cache.end() && it->second.refs > 0) it->second.refs--;
}

// Remove least recently used profiles beyond kServerProfiles that no running
// query holds (their tables are idle; msvg_profile_remove frees them).
void server_evict(msvg_server *s) {
  while (s->prof_cache.size() > kServerProfiles) {
    auto victim = s->prof_cache.end();
    for (auto it = s->prof_cache.begin(); it != s->prof_cache.end(); ++it)
      if (it->second.refs == 0 && (victim == s->prof_cache.end() || it->second.used < victim->second.used))
        victim = it;
    if (victim == s->prof_cache.end()) return;
    (void) msvg_profile_remove(s->ctx, victim->second.pid);
    s->prof_cache.erase(victim);
    std::lock_guard<std::mutex> lk(s->mu);
    s->cached = (int64_t) s->prof_cache.size();
  }
}

// The query's profile id on the device: from the cache, or added, its
// tables built on the slot's stream and waited for before the id is cached
// (a later query of the same profile may run on another slot's stream,
// which must not read a table still being built).  Takes a hold on it.
int server_profile(msvg_server *s, Slot &sl, ServerQuery &q, int32_t *ret_pid) {
  msvg_ctx *c = s->ctx;
  q.key = profile_key(q);
  auto it = s->prof_cache.find(q.key);
  if (it == s->prof_cache.end()) {
    q.desc.rbv = q.rbv.data();
    int32_t pid;
    int rc = msvg_profile_add(c, &q.desc, &pid);
    if (rc != MSVG_OK) return rc;
    int64_t cells0 = 0;
    rc = enqueue_search(c, sl.bufs, &pid, 1, s->db, s->F1, MSVG_MODE_BATCHED, false, false, &cells0, sl.stream,
                        s->sched, nullptr, true);
    if (rc == MSVG_OK && hipStreamSynchronize(sl.stream) != hipSuccess) rc = fail(c, MSVG_EDEVICE, "table build");
    if (rc != MSVG_OK) { (void) msvg_profile_remove(c, pid); return rc; }
    it = s->prof_cache.emplace(q.key, msvg_server::CachedProfile{pid, 0, 0}).first;
  }
  it->second.refs++;
  it->second.used = ++s->use_clock;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->cached = (int64_t) s->prof_cache.size();
  }
  *ret_pid = it->second.pid;
  return MSVG_OK;
}

// One search over the resident database for the slot's group of queries
// (one profile each; batched mode, so each lane-group size's profiles share
// one persistent launch), then each query's pass list compacted on the GPU.
int server_start_group(msvg_server *s, Slot &sl) {
  msvg_ctx *c = s->ctx;
  const Db &d = *c->dbs[s->db];
  const int k = (int) sl.group.size();
  const int64_t n = std::max<int64_t>(d.n, 1);
  std::vector<int32_t> pids((size_t) k);
  for (int j = 0; j < k; ++j) {
    const int rc = server_profile(s, sl, *sl.group[(size_t) j], &pids[(size_t) j]);
    if (rc != MSVG_OK) return rc;
  }
  int64_t cells = 0;
  int rc = enqueue_search(c, sl.bufs, pids.data(), k, s->db, s->F1, MSVG_MODE_BATCHED, false, false, &cells,
                          sl.stream, s->sched);
  if (rc != MSVG_OK) return rc;
  HIPCHK(c, sl.d_ccount.ensure((size_t) k * sizeof(int32_t)));
  HIPCHK(c, sl.d_cidx.ensure((size_t) k * n * sizeof(int32_t)));
  HIPCHK(c, sl.d_cusc.ensure((size_t) k * n * sizeof(float)));
  HIPCHK(c, msvk_fill_zero(sl.d_ccount.as<int32_t>(), k, sl.stream));
  for (int j = 0; j < k; ++j) {
    ServerQuery &q = *sl.group[(size_t) j];
    q.cells = (int64_t) q.desc.M * d.total_res;
    q.batch = k;
    HIPCHK(c, msvk_compact(sl.bufs.d_pass.as<uint8_t>() + (size_t) j * d.n, sl.bufs.d_usc.as<float>() + (size_t) j * d.n,
                           d.n, sl.d_ccount.as<int32_t>() + j, sl.d_cidx.as<int32_t>() + (size_t) j * n,
                           sl.d_cusc.as<float>() + (size_t) j * n, (int32_t) n, sl.stream));
  }
  HIPCHK(c, hipMemcpyAsync(sl.h_count, sl.d_ccount.p, (size_t) k * sizeof(int32_t), hipMemcpyDeviceToHost,
                           sl.stream));
  HIPCHK(c, hipEventRecord(sl.done, sl.stream));
  return MSVG_OK;
}

int server_finish_group(msvg_server *s, Slot &sl) {
  msvg_ctx *c = s->ctx;
  const int k = (int) sl.group.size();
  const int64_t n = std::max<int64_t>(c->dbs[s->db]->n, 1);
  // The byte-MSV queue (counters[k]) must not have overflowed; a search
  // that overflowed reports an error.
  int32_t n_slow = 0;
  HIPCHK(c, hipMemcpy(&n_slow, sl.bufs.d_counters.as<int32_t>() + k, sizeof(int32_t), hipMemcpyDeviceToHost));
  if (n_slow > sl.bufs.slow_cap) return fail(c, MSVG_EMEM, "deferred-pair queue overflow");
  for (int j = 0; j < k; ++j) {
    ServerQuery &q = *sl.group[(size_t) j];
    const int32_t m = sl.h_count[j];
    q.idx.resize((size_t) m);
    q.usc.resize((size_t) m);
    if (m > 0) {
      HIPCHK(c, hipMemcpy(q.idx.data(), sl.d_cidx.as<int32_t>() + (size_t) j * n, (size_t) m * sizeof(int32_t),
                          hipMemcpyDeviceToHost));
      HIPCHK(c, hipMemcpy(q.usc.data(), sl.d_cusc.as<float>() + (size_t) j * n, (size_t) m * sizeof(float),
                          hipMemcpyDeviceToHost));
    }
  }
  return MSVG_OK;             // the waiting caller orders the hits (msvg_server_wait)
}

// The slot's queries are over (rc: how): results released to the waiters.
void server_retire(msvg_server *s, Slot &sl, int rc) {
  for (ServerQuery *q : sl.group) server_release(s, *q);
  server_evict(s);
  std::lock_guard<std::mutex> lk(s->mu);
  const double t = s->now();
  for (ServerQuery *q : sl.group) {
    q->status = rc;
    q->state = 2;
    q->t_done = t;
    s->lds_in_use -= q->lds;
    s->dev_in_use -= q->dev;
    s->st.completed++;
    s->st.running--;
  }
  sl.group.clear();
  if (s->st.running == 0 && s->busy_since >= 0) { s->st.busy_s += t - s->busy_since; s->busy_since = -1; }
  s->cv.notify_all();
}

void server_loop(msvg_server *s) {
  msvg_ctx *c = s->ctx;
  (void) hipSetDevice(c->device);
  const Db &d = *c->dbs[s->db];
  for (;;) {
    bool progressed = false;
    // Completions.
    for (auto &slp : s->slots) {
      Slot &sl = *slp;
      if (sl.group.empty() || hipEventQuery(sl.done) != hipSuccess) continue;
      server_retire(s, sl, server_finish_group(s, sl));
      progressed = true;
    }
    // Admissions: a free slot, the HBM budget and the LDS budget (all running
    // queries' workgroups must be able to share a CU) decide.  A free slot
    // takes the first waiting query and, in batched serving, those behind it
    // up to kServerBatch: the queries that arrived while the slots were busy
    // go to the GPU together, in one search.
    for (;;) {
      Slot *free_slot = nullptr;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping || s->pending.empty()) break;
        for (auto &slp : s->slots)
          if (slp->group.empty()) { free_slot = slp.get(); break; }
        if (!free_slot) break;
        while (!s->pending.empty() && (int) free_slot->group.size() < s->batch_max) {
          ServerQuery *cand = s->pending.front();
          cand->lds = query_lds(c, d, cand->desc.M);
          cand->dev = (size_t) d.n * 16 + (size_t) cand->desc.M * 29 * 8;
          // Shared scheduling: the running queries' workgroups must fit a CU
          // together.  Persistent launches queue behind each other anyway.
          const bool first = s->st.running == 0;
          const bool lds_ok = s->sched != 2 || first || s->lds_in_use + cand->lds <= (size_t) kLdsBytes;
          const bool hbm_ok = s->hbm_budget <= 0 || first || (int64_t) (s->dev_in_use + cand->dev) <= s->hbm_budget;
          if (!lds_ok || !hbm_ok) break;
          s->pending.pop_front();
          cand->state = 1;
          cand->slot = (int) (free_slot - s->slots[0].get());
          cand->t_admit = s->now();
          s->lds_in_use += cand->lds;
          s->dev_in_use += cand->dev;
          s->st.admitted++;
          s->st.pending--;
          s->st.running++;
          if (s->busy_since < 0) s->busy_since = cand->t_admit;
          free_slot->group.push_back(cand);
        }
        if (free_slot->group.empty()) break;
        s->st.batches++;
      }
      const int rc = server_start_group(s, *free_slot);
      if (rc != MSVG_OK) server_retire(s, *free_slot, rc);
      progressed = true;
    }
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->stopping && s->st.running == 0) break;
      if (!progressed) s->cv.wait_for(lk, std::chrono::microseconds(s->st.running ? 50 : 2000));
    }
  }
}

}  // namespace

extern "C" {

int msvg_server_start(msvg_ctx *c, int32_t db_id, int32_t slots, double F1, int64_t hbm_budget, int32_t flags,
                      msvg_server **ret) {
  if (!c || !ret) return MSVG_EINVAL;
  if (flags != MSVG_SERVE_PERSISTENT && flags != MSVG_SERVE_SHARED && flags != MSVG_SERVE_BATCHED)
    return fail(c, MSVG_EINVAL, "unknown flags");
  if (!c->dbs.count(db_id)) return fail(c, MSVG_EINVAL, "unknown database id");
  if (slots < 1 || slots > 16) return fail(c, MSVG_EINVAL, "slots must be 1..16");
  (void) hipSetDevice(c->device);
  auto s = std::make_unique<msvg_server>();
  s->ctx = c;
  s->db = db_id;
  s->F1 = F1;
  s->hbm_budget = hbm_budget;
  s->sched = (flags == MSVG_SERVE_SHARED) ? 2 : 0;
  s->batch_max = (flags == MSVG_SERVE_BATCHED) ? kServerBatch : 1;
  for (int i = 0; i < slots; ++i) {
    auto sl = std::make_unique<Slot>();
    HIPCHK(c, hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
    HIPCHK(c, hipHostMalloc((void **) &sl->h_count, (size_t) s->batch_max * sizeof(int32_t), hipHostMallocDefault));
    s->slots.push_back(std::move(sl));
  }
  s->t0 = std::chrono::steady_clock::now();
  msvg_server *raw = s.get();
  s->th = std::thread(server_loop, raw);
  *ret = s.release();
  return MSVG_OK;
}

int msvg_server_submit(msvg_server *s, const msvg_oprofile *om, int64_t *ret_ticket) {
  if (!s || !om || !om->rbv || !ret_ticket) return MSVG_EINVAL;
  if (om->M < 1 || om->M > 16384 || om->Kp < 1 || om->Kp > 255 || (int) om->bias_b + (int) om->base_b >= 255)
    return MSVG_EINVAL;
  auto q = std::make_unique<ServerQuery>();
  q->desc = *om;
  q->rbv.assign(om->rbv, om->rbv + (size_t) om->M * om->Kp);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) return MSVG_EINVAL;
  q->ticket = s->next_ticket++;
  q->t_submit = s->now();
  *ret_ticket = q->ticket;
  s->pending.push_back(q.get());
  s->queries[q->ticket] = std::move(q);
  s->st.submitted++;
  s->st.pending++;
  s->st.max_pending = std::max<int64_t>(s->st.max_pending, (int64_t) s->pending.size());
  s->cv.notify_all();
  return MSVG_OK;
}

int msvg_server_wait(msvg_server *s, int64_t ticket, double timeout_s, msvg_query_result *out, int32_t *idx,
                     float *usc, int64_t cap) {
  if (!s || !out) return MSVG_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->queries.find(ticket);
  if (it == s->queries.end()) return MSVG_EINVAL;
  ServerQuery *q = it->second.get();
  auto ready = [&] { return q->state == 2; };
  if (timeout_s < 0) s->cv.wait(lk, ready);
  else if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    out->ticket = ticket;
    out->t_submit = q->t_submit;
    out->t_admit = q->t_admit;
    out->t_done = -1;
    return MSVG_ENORESULT;
  }
  std::unique_ptr<ServerQuery> own = std::move(it->second);
  s->queries.erase(it);
  lk.unlock();
  out->ticket = ticket;
  out->status = own->status;
  out->M = own->desc.M;
  out->n_pass = (int64_t) own->idx.size();
  out->cells = own->cells;
  out->t_submit = own->t_submit;
  out->t_admit = own->t_admit;
  out->t_done = own->t_done;
  out->batch = own->batch;
  if (idx && cap > 0) {
    const int64_t k = (int64_t) own->idx.size();
    std::vector<int32_t> ord(k);
    for (int64_t i = 0; i < k; ++i) ord[i] = (int32_t) i;
    std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return own->idx[a] < own->idx[b]; });
    const int64_t m = std::min<int64_t>(cap, k);
    for (int64_t i = 0; i < m; ++i) {
      idx[i] = own->idx[ord[i]];
      if (usc) usc[i] = own->usc[ord[i]];
    }
  }
  return MSVG_OK;
}

int msvg_server_stats(msvg_server *s, msvg_server_info *out) {
  if (!s || !out) return MSVG_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->uptime_s = s->now();
  out->cached_profiles = s->cached;
  if (s->busy_since >= 0) out->busy_s += out->uptime_s - s->busy_since;
  return MSVG_OK;
}

int msvg_server_stop(msvg_server *s) {
  if (!s) return MSVG_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    for (ServerQuery *q : s->pending) { q->state = 2; q->status = MSVG_ENORESULT; }
    s->st.pending = 0;
    s->pending.clear();
    s->cv.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  msvg_ctx *c = s->ctx;
  for (auto &sl : s->slots) {
    (void) hipStreamSynchronize(sl->stream);
    (void) hipEventDestroy(sl->done);
    (void) hipStreamDestroy(sl->stream);
    (void) hipHostFree(sl->h_count);
  }
  for (auto &kv : s->prof_cache) (void) msvg_profile_remove(c, kv.second.pid);
  delete s;
  return MSVG_OK;
}

}  // extern "C"
