// msv_device.h -- data layouts shared by the HIP kernels and the host runtime.
//
// Residues in HBM: one byte per residue, each target padded to a multiple of
// 16 bytes with the null row index, targets stored in descending-length order.
// Residue bytes are remapped at upload time to dense row indices 0..R-1 (the
// codes that occur in the database); row R is the null row (emission -inf).
//
// Fast-stage emission table (one per profile x database x lane group size G):
//   uint4 T[(R + 1) * QB * 64], QB = ceil(Q / 4)
//   entry (x, jb, lane) holds 4 dwords = registers j = 4*jb .. 4*jb+3;
//   each dword packs two f16 emissions for the two model slots of that lane:
//     slot z = 2*(lane % G) (low half) and 2*(lane % G) + 1 (high half),
//     model position k = j + z * Q + 1 (striped, as the reference stripes over
//     16 bytes, impl_sse/p7_oprofile.c:625-630, here over 2G slots).
//   value = (bias_b - rbv[x][k]) * 2^-11, or -1.0 for padding k > M and for the
//   null row.  The f16 domain stores W - B0 scaled by 2^-11, so the begin score
//   is 0.0 and the clamp bit of v_pk_add_f16 is the floor at the begin score.
//
// Slow-stage cost table (per profile x database): same striping with G = 64,
// Qs = ceil(M / 128), uint32 C[(R + 1) * Qs * 64], dword = (cost_lo, cost_hi)
// as two u16 holding rbv bytes (255 for padding and for the null row).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

#define MSVG_NULLCOST 255

// Per-pair result word written by the fast and slow stages:
//   bits 0..7 xJ, bits 8..9 code.  A deferred pair (RES_SLOW) carries bit 0
//   = 1 when only the J state is in doubt (the J-aware f16 stage resolves it)
//   and 0 when the exact sub-B0 maximum is needed (byte stage).
enum { RES_OK = 0, RES_ERANGE = 1, RES_SLOW = 2 };

struct SlowItem {          // one pair that needs the exact byte MSV
  int32_t prof;            // index into the search's profile array
  int32_t seq;             // sorted target index
};

struct ProfDev {           // per-profile constants seen by the kernels
  const uint4 *ftab;       // fast table (global), may be null if not built
  const uint32_t *stab;    // slow table (global)
  int32_t M, Q, QB, Qs;
  int32_t tbm, tec, base, bias;
  int32_t slow_all;        // 1: route every pair to the exact byte MSV
  uint16_t *res;           // result words for this profile, indexed by sorted target
};

struct DbDev {
  const uint8_t *res;      // remapped, padded residues
  const int64_t *off;      // sorted target -> byte offset
  const int32_t *len;      // sorted target -> length
  const uint8_t *tjb;      // sorted target -> tjb_b(L)
  int64_t n;
  int32_t R;               // residue rows in use; R is the null row
};

// Kernel argument blocks (passed by value).
struct FastArgs {
  ProfDev prof;
  DbDev db;
  const int32_t *chunk_start;  // dynamic mode: task ranges [chunk_start[c], chunk_start[c+1])
  int32_t nchunks;
  int32_t *counter;            // dynamic mode: next chunk (zeroed before the launch)
  const int32_t *order;        // static mode: task t -> sorted target index
  int32_t ntasks;              // static mode: number of tasks
  int32_t mode;                // 0 dynamic chunks, 1 static round-robin, 2 one chunk per workgroup
  int32_t prof_idx;
  SlowItem *slow;
  int32_t *slow_count;
  int32_t slow_cap;
  SlowItem *jslow;             // this profile's J-aware queue (null: use the byte queue)
  int32_t *jslow_count;
  int32_t jslow_cap;
};

// Batched J-free pass over many profiles in one persistent launch per lane
// group size G (msv_batch<G>).  Each profile has its own queue of target
// chunks (the database's guided chunk list for G, one counter per profile);
// a workgroup stages one profile's table, its waves drain that queue, and
// when it is empty the workgroup moves to the next profile with chunks left.
//
// Batched-pass tables (per profile x database x G), f16 emission pairs as in
// the fast-stage table (value (bias_b - rbv[x][k]) * 2^-11, -1.0 for padding
// and for the null row):
//   G = 64:      the fast-stage layout above (entry (x, jb, lane)).
//   G = 8, 16:   uint4 B[QB][32][16]: entry (jb, x, column), column = lane % 16
//                (slots of lane % G); 32 residue rows per register block.
//   G = 32:      uint4 B[QB][32][32]: entry (jb, x, lane % 32).
// With 32 rows per block the byte address of (jb, x, column) is
// jb * 32 * TW * 16 + x * TW * 16 + column * 16, so one v_perm_b32 builds the
// row address from a residue byte (x * 256 for TW = 16; (2x + column / 16) *
// 256 for TW = 32) and jb goes into the read's immediate offset.  Needs R + 1
// <= 32 and, for G = 32, QB <= 4.
struct BatProf {
  const uint4 *tab;            // batched-pass table
  ProfDev pd;                  // constants; pd.res receives the result words; pd.Q registers per lane
  int32_t *counter;            // next chunk of this profile's queue (zeroed before the launch)
  SlowItem *jslow;             // this profile's J-aware queue (null: deferred pairs go to the byte queue)
  int32_t *jslow_count;
  int32_t prof_idx;            // index into the search's profile array
  int32_t tab_u4;              // table size in uint4 entries
};

// Grouped scan (msv_scan_group<G, QB>): the J-free pass of many profiles,
// one per workgroup (b % np; the (b / np)-th of its splits), against a batch
// of queries.  Every profile of a launch has QB register blocks; pd.Q is
// its own register count and pd.ftab its table over the whole alphabet, in
// the batched pass's layout for G (above), over codes 0..Kp-1 and the
// floor rows after them.  The workgroup stages rows rowcode[0..db.R] of it
// (the codes the queries hold, then a floor row as the null row R) as LDS
// rows 0..R, each register block keeping its stride of 32 rows.
struct ScanGroupArgs {
  const ProfDev *profs;
  int32_t np;
  DbDev db;                    // the queries, longest first, residues as LDS rows (db.R = null row)
  const uint8_t *rowcode;      // LDS row r holds table row rowcode[r] (r = 0..db.R)
  const int32_t *soff;         // wave slot w = split * waves + wave scores the tasks
  const int32_t *stask;        //   stask[soff[w] .. soff[w + 1]) (task t = queries t*64/G ..)
  uint16_t *out;               // word of (profile b % np, query s) at out[s * stride + b % np]
  int64_t stride;
  int32_t prio;                // 1: the waves raise their issue priority (a launch whose few
                               // workgroups' row chains bound the batch, beside busier launches)
};

struct BatArgs {
  const BatProf *profs;
  int32_t np;
  DbDev db;
  const int32_t *chunk_start;  // chunk c = tasks [chunk_start[c], chunk_start[c+1]) (task = 64/G targets)
  int32_t nchunks;
  const int32_t *first;        // profile each workgroup starts on
  SlowItem *slow;
  int32_t *slow_count;
  int32_t slow_cap;
  int32_t jslow_cap;
  int32_t slot_u4;             // LDS uint4 index of the workgroup's broadcast word (past the largest table)
};

// J-aware stage: the MSV recurrence with per-row E/J/B updates in the f16
// domain of the fast stage, one target per wave, for one profile's queue.
struct JArgs {
  ProfDev prof;
  DbDev db;
  const uint4 *tab;            // the profile's fast table for 64-lane groups
  const SlowItem *items;
  const int32_t *count;
  int32_t cap;
};

// msv_jrow_many: the J-aware pairs of many profiles in one launch.
// Workgroup b works on profile b % np; its waves take that profile's pairs
// at a fixed stride (no device counter).
struct JProf {
  ProfDev prof;
  const uint4 *tab;            // the profile's fast table for 64-lane groups, q registers
  const SlowItem *items;
  const int32_t *count;        // pairs queued (by the J-free pass)
  int32_t cap, q;
};

struct JManyArgs {
  DbDev db;
  const JProf *profs;
  int32_t np;
  int32_t wpp;                 // workgroups per profile (grid = np * wpp)
};

struct OneArgs {               // one target, one workgroup (the per-call drop-in)
  ProfDev prof;
  const uint4 *tab;            // 64-lane fast table over residue codes 0..R-1, null row R
  const uint8_t *res;          // the target's codes, padded to 16 bytes with R
  int32_t L, tjb, R;
  uint32_t *out;               // result word
};

struct ScanProf {              // one profile of a scan set (many profiles, one query)
  const uint4 *tab;            // its 64-lane table over residue codes 0..R-1, null row R
  ProfDev prof;
};

struct ScanArgs {              // one workgroup per profile, its waves over the queries
  const ScanProf *profs;
  const uint8_t *res;          // the queries' codes, each padded to 16 bytes with R
  const int64_t *qoff;         // query q starts at res + qoff[q]
  const int32_t *qlen;
  const uint8_t *qtjb;         // tjb_b(L) per query
  int32_t nq, R;
  int32_t stride;              // result of (profile b, query q) at out[q * stride + b]
  uint32_t *out;
  const int32_t *wq_off;       // non-null: workgroup b scores only the queries wq_idx[wq_off[b] ..
  const int32_t *wq_idx;       // wq_off[b+1]), the result of item k at out[k]
  const uint8_t *rowcode;      // LDS row r holds table row rowcode[r] (r = 0..R; R is the null row)
};

struct SlowArgs {
  const ProfDev *profs;
  DbDev db;
  const SlowItem *items;
  const int32_t *count;
  int32_t cap;
};

struct FinArgs {
  const ProfDev *profs;
  DbDev db;
  const int64_t *sorted_pos;   // caller target t -> sorted index
  const int32_t *lidx;         // sorted index -> distinct-length index
  const float *usc_tab;        // [np][511]: usc for xJ - tjb + 255
  const uint16_t *thr;         // [np][nL]: smallest passing xJ (256 = none)
  int32_t nL;
  float *usc;                  // [np][n], may be null
  uint8_t *status;             // [np][n], may be null
  uint8_t *pass;               // [np][n], may be null
};

// Column-segmented fast pass for long models: one launch per model segment
// of 128*Q positions; the value leaving the segment's last position in row i
// is stored in carry[off(s) + i] (f16 bits) and enters the next segment's
// first position in row i+1.  partial[s] keeps the running maximum (units).
struct SegArgs {
  ProfDev prof;            // Q/QB are the segment's
  DbDev db;
  const uint4 *tab;        // segment table, fast-table layout with G = 64
  uint16_t *carry;         // per target, indexed like the residues
  uint16_t *partial;       // per sorted target
  const int64_t *carry_base;
  int32_t first, last;
  int32_t ntasks;
  int32_t prof_idx;
  SlowItem *slow;
  int32_t *slow_count;
  int32_t slow_cap;
  SlowItem *jslow;             // the profile's J-aware queue, or null
  int32_t *jslow_count;
  int32_t jslow_cap;
};

// ---------------------------------------------------------------------------
// ViterbiFilter stage (impl_sse/vitfilter.c:84-250) for listed pairs.
//
// W waves (64*W lanes, 128*W int16 slots) own one target; register j of the
// lane with index c = 64*w + lane holds slots 2c (low half) and 2c+1 (high
// half), model position k = j + s*Q + 1 for slot s (Q = ceil(M / (128*W))).
// Tables, one u32 (two int16) per (row, j, c):
//   transitions  T[(t * Q + j) * 64W + c], t = BM, MM, IM, DM, MD, MI, II, DD,
//                the values of the reference's striped vectors at k (rotated
//                BM/MM/IM/DM: into k from k-1), -32768 past the model;
//   emissions    E[(x * Q + j) * 64W + c] for database row x (R+1 rows; the
//                null row is never read: rows run over the real residues).
// Result per item: (status << 16) | (uint16) xC, status 0 or 16 (eslERANGE);
// the host turns xC into the score with the reference's float expression.
struct VitProfDev {            // one profile of a vit_band launch
  const uint32_t *etab;        // emissions, E layout above
  const uint32_t *ttab;        // transitions, T layout above
  int32_t e_loop, e_move, base_w, ddbound_w;
  int32_t end;                 // its items end at this position of the item arrays
};

struct VitArgs {
  const uint32_t *etab;
  const uint32_t *ttab;
  DbDev db;
  const int64_t *items;        // sorted target index per item
  const int16_t *move;         // xw[N|C|J][MOVE] for the item's length
  int32_t n;
  int32_t *out;
  int32_t e_loop, e_move, base_w, ddbound_w;
  const int32_t *uoff;         // slot k (a wave, or a workgroup of a multi-wave
                               // shape) runs units [uoff[k], uoff[k+1]); a unit
                               // is one wave's targets
  // vit_band only (several profiles per launch): workgroup b works on
  // profile profs[wg_prof[b]]; unit u's targets are the items at positions
  // ustart[u] + 0 .. 64/G - 1 below that profile's end.  etab, ttab, n and
  // the scalars above are not read.
  const VitProfDev *profs;
  const int32_t *wg_prof;
  const int32_t *ustart;
};

// vit_seg: one column segment of a long model (its positions k0 + 1 ..
// k0 + 128 Q, tables in the 64-lane layout of vit_band), one target per wave.
// Per item row the segment reads what the segment before it left (the M, I
// and D values at its last position, and the D value entering the next
// position from it) and writes its own in their place; the items' largest
// row maximum and whether any row could raise J (or overflow) carry over in
// partial[].  The last segment writes the result word, or VIT_DEFER for an
// item some row of which could change B: such items are scored again by the
// exact multi-wave kernel.
#define VIT_DEFER (1 << 24)
struct VitSegArgs {
  DbDev db;
  const uint32_t *etab, *ttab;
  int32_t e_loop, e_move, base_w, ddbound_w;
  const int64_t *items;        // item -> sorted target index
  const int16_t *move;         // item -> xw[N|C|J][MOVE] for its length
  const int64_t *coff;         // item -> its first carry row
  int16_t *cM, *cI, *cD, *cE;  // carries per item row
  int32_t *partial;            // item -> largest row maximum (low 16 bits) | could-fire (bit 16)
  int32_t *out;                // item -> result word (last segment)
  const int32_t *uoff;         // wave slot k runs items [uoff[k], uoff[k+1])
  int32_t first, last;
};
