// msv_host.cpp -- host-side profile preparation and workload generation
// (declared in include/msv_prep.h).  Every routine follows the float/double
// evaluation order of the HMMER/Easel function it mirrors, because the byte
// profile, tjb_b and the pass rule are defined by those roundings.  Build
// with -ffp-contract=off and without fast-math.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/msv_prep.h"
#include "msv_host.h"

namespace {

constexpr double kLog2 = 0.69314718055994529;   // eslCONST_LOG2, easel/easel.h:258

uint32_t jenkins_mix3(uint32_t a, uint32_t b, uint32_t c) {   // esl_random.c:383-398
  a -= b; a -= c; a ^= (c >> 13);
  b -= c; b -= a; b ^= (a << 8);
  c -= a; c -= b; c ^= (b >> 13);
  a -= b; a -= c; a ^= (c >> 12);
  b -= c; b -= a; b ^= (a << 16);
  c -= a; c -= b; c ^= (b >> 5);
  a -= b; a -= c; a ^= (c >> 3);
  b -= c; b -= a; b ^= (a << 10);
  c -= a; c -= b; c ^= (b >> 15);
  return c;
}

void mt_fill(msvh_rng *r) {                      // esl_random.c:332-354 (MT19937 refill)
  static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
  uint32_t y;
  int z;
  for (z = 0; z < 227; z++) {
    y = (r->mt[z] & 0x80000000u) | (r->mt[z + 1] & 0x7fffffffu);
    r->mt[z] = r->mt[z + 397] ^ (y >> 1) ^ mag01[y & 1u];
  }
  for (; z < 623; z++) {
    y = (r->mt[z] & 0x80000000u) | (r->mt[z + 1] & 0x7fffffffu);
    r->mt[z] = r->mt[z - 227] ^ (y >> 1) ^ mag01[y & 1u];
  }
  y = (r->mt[623] & 0x80000000u) | (r->mt[0] & 0x7fffffffu);
  r->mt[623] = r->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
  r->mti = 0;
}

uint32_t rng_u32(msvh_rng *r) {
  if (r->fast) {                                 // knuth(), esl_random.c:278-284
    r->x = r->x * 69069u + 1u;
    return r->x;
  }
  if (r->mti >= 624) mt_fill(r);                 // mersenne_twister(), :297-308
  uint32_t x = r->mt[r->mti++];
  x ^= (x >> 11);
  x ^= (x << 7) & 0x9d2c5680u;
  x ^= (x << 15) & 0xefc60000u;
  x ^= (x >> 18);
  return x;
}

double uniform_positive(msvh_rng *r) {           // esl_rnd_UniformPositive, :449-454
  double x;
  do { x = msvh_random(r); } while (x == 0.0);
  return x;
}

// esl_dirichlet_FSampleUniform (esl_dirichlet.c:964-971): Gamma(1) deviates
// (esl_rnd_Gamma -> gamma_integer with a = 1, esl_random.c:607-617, :657-671)
// normalized by esl_vec_FNorm (esl_vectorops.c:842-850).
void dirichlet_uniform(msvh_rng *r, int K, float *p) {
  for (int x = 0; x < K; x++) {
    double U = 1.;
    U *= uniform_positive(r);
    p[x] = (float) (-log(U));
  }
  float sum = 0.;
  for (int x = 0; x < K; x++) sum += p[x];
  if (sum != 0.0) for (int x = 0; x < K; x++) p[x] /= sum;
  else            for (int x = 0; x < K; x++) p[x] = 1. / (float) K;
}

const float kAmino[20] = {                       // p7_AminoFrequencies, hmmer.c:163-186
  0.0787945f, 0.0151600f, 0.0535222f, 0.0668298f, 0.0397062f, 0.0695071f, 0.0229198f,
  0.0590092f, 0.0594422f, 0.0963728f, 0.0237718f, 0.0414386f, 0.0482904f, 0.0395639f,
  0.0540978f, 0.0683364f, 0.0540687f, 0.0673417f, 0.0114135f, 0.0304133f};

// The digital alphabets HMMER profiles use (esl_alphabet.c:180-300):
// K canonical residues, Kp codes in all, background frequencies as
// p7_bg_Create sets them (p7_bg.c:55-82: p7_AminoFrequencies for amino,
// uniform 1/K otherwise), and for each degenerate code K+1 .. Kp-3 the set
// of residues it stands for.
struct Alphabet {
  int K = 0, Kp = 0;
  float f[20] = {};
  uint8_t degen[MSVH_KP][20] = {};
};

bool alphabet_of(int type, Alphabet &a) {
  auto set = [&a](int x, const char *members, const char *order) {
    for (const char *m = members; *m; ++m) a.degen[x][std::strchr(order, *m) - order] = 1;
  };
  if (type == MSVH_AMINO) {
    static const char order[] = "ACDEFGHIKLMNPQRSTVWY";
    a.K = 20; a.Kp = 29;
    std::memcpy(a.f, kAmino, sizeof kAmino);
    set(21, "ND", order); set(22, "IL", order); set(23, "QE", order);    // B, J, Z
    set(24, "K", order);  set(25, "C", order);                            // O, U
    set(26, order, order);                                                // X
    return true;
  }
  if (type == MSVH_DNA || type == MSVH_RNA) {
    static const char order[] = "ACGT";                   // U takes T's place in RNA
    a.K = 4; a.Kp = 18;
    for (int x = 0; x < 4; x++) a.f[x] = 1. / (float) 4;
    static const char *const d[11] = {"AG", "CT", "AC", "GT", "CG", "AT", "ACT", "CGT", "ACG", "AGT", "ACGT"};
    for (int i = 0; i < 11; i++) set(5 + i, d[i], order);                // R Y M K S W H B V D N
    return true;
  }
  return false;
}

// unbiased_byteify (impl_sse/p7_oprofile.c:512-519).
uint8_t unbiased_byteify(float scale_b, float sc) {
  sc = -1.0f * roundf(scale_b * sc);
  return (sc > 255.) ? 255 : (uint8_t) sc;
}

// biased_byteify (impl_sse/p7_oprofile.c:495-503).  The reference's
// (uint8_t) cast of a negative float is carried out on x86-64 as a
// conversion to a 32-bit integer whose low byte is kept; the same is written
// out here so the result does not depend on this compiler's choice.
uint8_t biased_byteify(float scale_b, int bias_b, float sc) {
  sc = -1.0f * roundf(scale_b * sc);
  if (sc > 255 - bias_b) return 255;
  const uint8_t low = (uint8_t) (uint32_t) (int32_t) sc;
  return (uint8_t) (low + bias_b);
}

}  // namespace

extern "C" {

void msvh_rng_init(msvh_rng *r, uint32_t seed, int fast) {   // esl_randomness_Init, :197-214
  std::memset(r, 0, sizeof(*r));
  r->fast = fast ? 1 : 0;
  if (seed == 0) seed = 42;
  r->seed = seed;
  if (!r->fast) {
    r->mt[0] = seed;                                           // mersenne_seed_table, :315-325
    for (int z = 1; z < 624; z++) r->mt[z] = 69069u * r->mt[z - 1];
    mt_fill(r);
  } else {
    r->x = jenkins_mix3(seed, 87654321u, 12345678u);
    if (r->x == 0) r->x = 42;
  }
}

double msvh_random(msvh_rng *r) { return ((double) rng_u32(r) / 4294967296.0); }   // esl_random, :271-276

int msvh_fchoose(msvh_rng *r, const float *p, int N) {     // esl_rnd_FChoose, :718-738
  double norm = 0.0, sum = 0.0;
  const double roll = msvh_random(r);
  for (int i = 0; i < N; i++) norm += p[i];
  for (int i = 0; i < N; i++) {
    sum += p[i];
    if (roll < (sum / norm)) return i;
  }
  return N - 1;   // unreachable for a normalized p (the reference aborts here)
}

void msvh_background(float *f) { std::memcpy(f, kAmino, sizeof(kAmino)); }

int msvh_hmm_sample(msvh_rng *r, int M, float *mat, float *ins, float *t) {   // p7_hmm_Sample
  if (M < 1) return 11;
  std::memset(mat, 0, sizeof(float) * (M + 1) * 20);
  std::memset(ins, 0, sizeof(float) * (M + 1) * 20);
  std::memset(t, 0, sizeof(float) * (M + 1) * 7);
  t[5] = 1.0;                           // D_0 does not exist: t[0][DM] = 1 (p7_hmm.c:163)
  for (int k = 0; k <= M; k++) {
    if (k > 0) dirichlet_uniform(r, 20, mat + k * 20);
    dirichlet_uniform(r, 20, ins + k * 20);
    dirichlet_uniform(r, 3, t + k * 7);
    dirichlet_uniform(r, 2, t + k * 7 + 3);
    if (k > 0) dirichlet_uniform(r, 2, t + k * 7 + 5);
  }
  dirichlet_uniform(r, 2, t + M * 7);   // node M: TMM, TMI only
  t[M * 7 + 2] = 0.;                    // MD
  t[M * 7 + 5] = 1.0;                   // DM
  t[M * 7 + 6] = 0.0;                   // DD
  return 0;
}

int msvh_profile_convert_abc(int alphatype, int M, const float *mat, int config_L, uint8_t *rbv, int *scalars,
                             float *scale, float *msc_out) {
  Alphabet ab;
  if (M < 1 || !alphabet_of(alphatype, ab)) return 11;
  const int K = ab.K, Kp = ab.Kp;
  const float *f = ab.f;
  // Match emission scores, modelconfig.c:138-150: log(p / f) for the
  // residues, esl_abc_FExpectScVec (esl_alphabet.c:1399-1413, :1507-1513)
  // for the degenerate codes, -inf for gap, nonresidue and missing data.
  std::vector<float> msc((size_t) Kp * M);
  float sc[MSVH_KP];
  for (int k = 1; k <= M; k++) {
    for (int x = 0; x < K; x++) sc[x] = log(mat[k * K + x] / f[x]);
    sc[K] = -INFINITY;
    sc[Kp - 2] = -INFINITY;
    sc[Kp - 1] = -INFINITY;
    for (int d = K + 1; d <= Kp - 3; d++) {
      float result = 0., denom = 0.;
      for (int i = 0; i < K; i++)
        if (ab.degen[d][i]) { result += sc[i] * f[i]; denom += f[i]; }
      result /= denom;
      sc[d] = result;
    }
    for (int x = 0; x < Kp; x++) msc[(size_t) x * M + (k - 1)] = sc[x];
  }
  // mf_conversion, p7_oprofile.c:602-640.  The maximum runs over the real
  // residues' match and insert scores; inserts are 0 for k < M.
  float max = 0.0;
  for (int x = 0; x < K; x++)
    for (int k = 1; k <= M; k++) max = std::max(max, msc[(size_t) x * M + (k - 1)]);
  const float scale_b = 3.0 / kLog2;
  const int base_b = 190;
  const int bias_b = unbiased_byteify(scale_b, -1.0 * max);
  for (int x = 0; x < Kp; x++)
    for (int k = 1; k <= M; k++)
      rbv[(size_t) x * M + (k - 1)] = biased_byteify(scale_b, bias_b, msc[(size_t) x * M + (k - 1)]);
  scalars[0] = unbiased_byteify(scale_b, logf(2.0f / ((float) M * (float) (M + 1))));   // tbm_b
  scalars[1] = unbiased_byteify(scale_b, logf(0.5f));                                    // tec_b
  scalars[2] = unbiased_byteify(scale_b, logf(3.0f / (float) (config_L + 3)));           // tjb_b
  scalars[3] = base_b;
  scalars[4] = bias_b;
  *scale = scale_b;
  if (msc_out) std::memcpy(msc_out, msc.data(), sizeof(float) * msc.size());
  return 0;
}

int msvh_profile_convert(int M, const float *mat, int config_L, uint8_t *rbv, int *scalars, float *scale,
                         float *msc_out) {
  return msvh_profile_convert_abc(MSVH_AMINO, M, mat, config_L, rbv, scalars, scale, msc_out);
}

namespace {
// wordify(), p7_oprofile.c:527-534
int16_t wordify(float scale_w, float sc) {
  sc = roundf(scale_w * sc);
  if (sc >= 32767.0) return 32767;
  else if (sc <= -32768.0) return -32768;
  else return (int16_t) sc;
}
}  // namespace

int msvh_vit_convert_abc(int alphatype, int M, const float *mat, const float *t, int16_t *rwv, int16_t *twv,
                         int *scalars, float *scale_w_out, float *nj_out) {
  Alphabet ab;
  if (M < 1 || !alphabet_of(alphatype, ab)) return 11;
  const int Kp = ab.Kp;
  std::vector<float> msc((size_t) Kp * M);
  std::vector<uint8_t> rbv((size_t) Kp * M);
  int sc5[5];
  float scale_b;
  int rc = msvh_profile_convert_abc(alphatype, M, mat, 100, rbv.data(), sc5, &scale_b, msc.data());
  if (rc != 0) return rc;
  enum { MM = 0, IM = 1, DM = 2, BM = 3, MD = 4, MI = 5, II = 6, DD = 7 };   // p7P_* (hmmer.h:222-231)
  enum { hMM = 0, hMI = 1, hMD = 2, hIM = 3, hII = 4, hDM = 5, hDD = 6 };   // p7H_*
  // Generic transition scores for nodes 0..M-1: node 0 has none but the
  // entry (p7_profile.c:84), local entry occ[k] / sum occ[i](M-i+1) with
  // p7_hmm_CalculateOccupancy (modelconfig.c:88-98, p7_hmm.c:1271-1281), the
  // rest log(t) (modelconfig.c:127-136).
  std::vector<float> tsc((size_t) M * 8, -INFINITY);
  std::vector<float> occ((size_t) M + 1);
  occ[0] = 0.;
  occ[1] = t[0 * 7 + hMI] + t[0 * 7 + hMM];
  for (int k = 2; k <= M; k++)
    occ[k] = occ[k - 1] * (t[(k - 1) * 7 + hMM] + t[(k - 1) * 7 + hMI]) + (1.0 - occ[k - 1]) * t[(k - 1) * 7 + hDM];
  float Z = 0.;
  for (int k = 1; k <= M; k++) Z += occ[k] * (float) (M - k + 1);
  for (int k = 1; k <= M; k++) tsc[(size_t) (k - 1) * 8 + BM] = log(occ[k] / Z);
  for (int k = 1; k < M; k++) {
    float *tp = &tsc[(size_t) k * 8];
    tp[MM] = log(t[k * 7 + hMM]);
    tp[MI] = log(t[k * 7 + hMI]);
    tp[MD] = log(t[k * 7 + hMD]);
    tp[IM] = log(t[k * 7 + hIM]);
    tp[II] = log(t[k * 7 + hII]);
    tp[DM] = log(t[k * 7 + hDM]);
    tp[DD] = log(t[k * 7 + hDD]);
  }
  auto TSC = [&](int k, int s) { return tsc[(size_t) k * 8 + s]; };
  // vf_conversion, p7_oprofile.c:653-740, in model order.
  const float scale_w = 500.0 / kLog2;                                // :676
  for (int x = 0; x < Kp; x++)                                        // :680-685
    for (int k = 1; k <= M; k++) rwv[(size_t) x * M + k - 1] = wordify(scale_w, msc[(size_t) x * M + k - 1]);
  // The seven per-node transitions (:688-707): BM, MM, IM, DM from node
  // k-1 (rotated), MD, MI, II from node k; capped at 0 (II at -1).
  const int from[7] = {BM, MM, IM, DM, MD, MI, II};
  for (int tt = 0; tt < 7; tt++) {
    const int maxval = (tt == 6) ? -1 : 0;
    for (int k = 1; k <= M; k++) {
      const int kb = (tt < 4) ? k - 1 : k;
      const int16_t val = (kb < M) ? wordify(scale_w, TSC(kb, from[tt])) : (int16_t) -32768;
      twv[(size_t) tt * M + k - 1] = (val <= maxval) ? val : (int16_t) maxval;
    }
  }
  for (int k = 1; k <= M; k++)                                        // DD, :710-715
    twv[(size_t) 7 * M + k - 1] = (k < M) ? wordify(scale_w, TSC(k, DD)) : (int16_t) -32768;
  const float e_sc = -kLog2;                                          // multihit E, modelconfig.c:116-117
  scalars[0] = wordify(scale_w, e_sc);                                // xw[E][LOOP], :724
  scalars[1] = wordify(scale_w, e_sc);                                // xw[E][MOVE], :725
  scalars[2] = 12000;                                                 // base_w, :677
  int16_t ddbound = -32768;                                           // :738-745
  for (int k = 2; k < M - 1; k++) {
    int ddtmp = (int) wordify(scale_w, TSC(k, DD));
    ddtmp += (int) wordify(scale_w, TSC(k + 1, DM));
    ddtmp -= (int) wordify(scale_w, TSC(k + 1, BM));
    ddbound = (int16_t) std::max((int) ddbound, ddtmp);
  }
  scalars[3] = ddbound;
  *scale_w_out = scale_w;
  *nj_out = 1.0f;                                                     // multihit, modelconfig.c:119
  return 0;
}

int msvh_vit_convert(int M, const float *mat, const float *t, int16_t *rwv, int16_t *twv, int *scalars,
                     float *scale_w_out, float *nj_out) {
  return msvh_vit_convert_abc(MSVH_AMINO, M, mat, t, rwv, twv, scalars, scale_w_out, nj_out);
}

double msvh_lambda_abc(int alphatype, int M, const float *mat) {
  // p7_Lambda (evalues.c:186-192) with p7_MeanMatchRelativeEntropy
  // (modelstats.c:81-98) and esl_vec_FRelEntropy (esl_vectorops.c:950-962).
  Alphabet ab;
  if (M < 1 || !alphabet_of(alphatype, ab)) return NAN;
  const int K = ab.K;
  double KL = 0.;
  for (int k = 1; k <= M; k++) {
    const float *p = mat + k * K;
    float kl = 0.;
    bool inf = false;
    for (int i = 0; i < K; i++)
      if (p[i] > 0.) {
        if (ab.f[i] == 0.) { inf = true; break; }
        kl += p[i] * log(p[i] / ab.f[i]);
      }
    const float rel = inf ? INFINITY : (float) (1.44269504 * kl);
    KL += rel;
  }
  KL /= (double) M;
  return kLog2 + 1.44 / ((double) M * KL);
}

double msvh_lambda(int M, const float *mat) { return msvh_lambda_abc(MSVH_AMINO, M, mat); }

int msvh_core_emit(msvh_rng *r, int M, const float *mat, const float *ins, const float *t, uint8_t *out, int cap) {
  // p7_CoreEmit, emit.c:71-148.  States B=0, M=1, I=2, D=3, E=4.
  enum { B, MS, IS, DS, ES };
  int k = 0, n = 0, st = B;
  while (st != ES) {
    switch (st) {
      case B:
      case MS: {
        const int c = msvh_fchoose(r, t + k * 7, 3);
        st = (c == 0) ? MS : (c == 1) ? IS : DS;
        break;
      }
      case IS: st = (msvh_fchoose(r, t + k * 7 + 3, 2) == 0) ? MS : IS; break;
      case DS: st = (msvh_fchoose(r, t + k * 7 + 5, 2) == 0) ? MS : DS; break;
    }
    if (st == MS || st == DS) k++;
    if (k == M + 1) {
      if (st == MS) { st = ES; k = 0; }
      else return -2;   // the reference raises eslECORRUPT here
    }
    int x = -1;
    if (st == MS) x = msvh_fchoose(r, mat + k * 20, 20);
    else if (st == IS) x = msvh_fchoose(r, ins + k * 20, 20);
    if (x >= 0) {
      if (n >= cap) return -1;
      out[n++] = (uint8_t) x;
    }
  }
  return n;
}

void msvh_iid(msvh_rng *r, int L, uint8_t *out) {          // esl_rsq_xfIID
  for (int i = 0; i < L; i++) out[i] = (uint8_t) msvh_fchoose(r, kAmino, 20);
}

int msvh_tjb(int L) { return msvh_tjb_impl(L); }
float msvh_nullsc(int L) { return msvh_nullsc_impl(L); }

double msvh_fit_mu(const float *usc, const int32_t *status, int N, int L, int base_b, float scale_b, double lambda) {
  const float maxsc = (255 - base_b) / scale_b;             // evalues.c:248
  const float nullsc = msvh_nullsc_impl(L);
  double esum = 0.;
  for (int i = 0; i < N; i++) {
    const float sc = (status[i] == 16) ? maxsc : usc[i];
    const double xv = (sc - nullsc) / kLog2;                // evalues.c:270
    esum += exp(-lambda * xv);                              // esl_gumbel.c:498-500
  }
  return -log(esum / N) / lambda;
}

void msvh_calibration_targets(int N, int L, uint8_t *out) {
  msvh_rng r;
  msvh_rng_init(&r, 42, 1);
  for (int i = 0; i < N; i++) msvh_iid(&r, L, out + (size_t) i * L);
}

int msvh_vit_move(float scale_w, float nj, int L) {
  const float pmove = (2.0f + nj) / ((float) L + 2.0f + nj);  // p7_oprofile.c:948
  float sc = logf(pmove);                                     // :956
  sc = roundf(scale_w * sc);                                  // wordify, :530-533
  if (sc >= 32767.0) return 32767;
  else if (sc <= -32768.0) return -32768;
  else return (int16_t) sc;
}

// How often p7_ViterbiFilter's lazy-F test (Dmax + ddbound_w > xB,
// vitfilter.c:199) holds, per row, on one pseudo-random target of `rows`
// residues (uniform over the K = 20 or 4 canonical codes): the recurrence of
// vitfilter.c:131-227 in model order, in int arithmetic clamped to int16.
// A planning estimate for the kernels (whether a row should vote on the test
// or always close the delete paths -- the results are the same either way),
// not a score.
double msvh_vit_lazyf_rate(const int16_t *rwv, const int16_t *twv, int M, int Kp, int e_loop, int base_w,
                           int ddbound_w, int move, int rows) {
  if (M < 1 || rows < 1) return 0.0;
  const int K = (Kp >= 26) ? 20 : 4;
  auto cl = [](int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); };
  auto T = [&](int t, int k) { return (int) twv[(size_t) t * M + k - 1]; };   // k = 1..M
  std::vector<int> Mv(M + 2, -32768), Iv(M + 2, -32768), Dv(M + 2, -32768), Mn(M + 2), In(M + 2), Dn(M + 2);
  const int xN = base_w;
  int xB = cl(xN + move), xJ = -32768, fired = 0;
  uint32_t r = 12345u;
  for (int i = 0; i < rows; ++i) {
    r = r * 1103515245u + 12345u;
    const int x = (int) ((r >> 16) % (uint32_t) K);
    int xE = -32768, Dmax = -32768;
    Mn[0] = In[0] = Dn[0] = -32768;
    for (int k = 1; k <= M; ++k) {
      int sv = cl(xB + T(0, k));
      sv = std::max(sv, cl(Mv[k - 1] + T(1, k)));
      sv = std::max(sv, cl(Iv[k - 1] + T(2, k)));
      sv = std::max(sv, cl(Dv[k - 1] + T(3, k)));
      sv = cl(sv + rwv[(size_t) x * M + k - 1]);
      Mn[k] = sv;
      xE = std::max(xE, sv);
      In[k] = std::max(cl(Mv[k] + T(5, k)), cl(Iv[k] + T(6, k)));
      Dn[k] = (k > 1) ? cl(Mn[k - 1] + T(4, k - 1)) : -32768;
      Dmax = std::max(Dmax, Dn[k]);
    }
    if (xE >= 32767) break;
    xJ = cl(std::max(xJ, xE + e_loop));
    xB = cl(std::max(xJ + move, xN + move));
    if (Dmax + ddbound_w > xB) {
      ++fired;
      for (int k = 2; k <= M; ++k) Dn[k] = std::max(Dn[k], cl(Dn[k - 1] + T(7, k - 1)));
    }
    Mv.swap(Mn); Iv.swap(In); Dv.swap(Dn);
  }
  return (double) fired / rows;
}

float msvh_vit_score(int xC, int move, int base_w, float scale_w) {
  if (xC <= -32768) return -INFINITY;                         // vitfilter.c:240, :247
  float sc = (float) xC + (float) move - (float) base_w;      // :242
  sc /= scale_w;                                              // :244
  sc -= 3.0;                                                  // :245
  return sc;
}

double msvh_fit_vmu(const float *sc, const int32_t *status, int N, int L, int base_w, float scale_w, double lambda) {
  const float maxsc = (32767.0 - base_w) / scale_w;           // evalues.c:321
  const float nullsc = msvh_nullsc_impl(L);
  double esum = 0.;
  for (int i = 0; i < N; i++) {
    const float s = (status[i] == 16) ? maxsc : sc[i];
    const double xv = (s - nullsc) / kLog2;                   // evalues.c:343
    esum += exp(-lambda * xv);                                // esl_gumbel.c:498-500
  }
  return -log(esum / N) / lambda;
}

void msvh_vit_calibration_targets(int N, int L, uint8_t *out) {
  msvh_rng r;
  msvh_rng_init(&r, 42, 1);
  std::vector<uint8_t> skip((size_t) L);
  for (int i = 0; i < N; i++) msvh_iid(&r, L, skip.data());   // p7_MSVMu's draws
  for (int i = 0; i < N; i++) msvh_iid(&r, L, out + (size_t) i * L);
}

int64_t msvh_gen_lengths(uint32_t seed, int64_t n, double median, double sigma, int lo, int hi, double tail_frac,
                         double tail_xm, double tail_alpha, int tail_max, int32_t *lens) {
  msvh_rng r;
  msvh_rng_init(&r, seed, 0);
  int64_t total = 0;
  const double two_pi = 6.283185307179586;
  for (int64_t i = 0; i < n; i++) {
    int L;
    if (tail_frac > 0.0 && msvh_random(&r) < tail_frac) {
      const double u = uniform_positive(&r);
      double x = tail_xm * pow(u, -1.0 / tail_alpha);
      if (x > tail_max) x = tail_max;
      L = (int) x;
    } else {
      const double u1 = uniform_positive(&r), u2 = msvh_random(&r);
      const double z = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
      double x = median * exp(sigma * z);
      L = (int) std::lround(x);
      if (L < lo) L = lo;
      if (L > hi) L = hi;
    }
    lens[i] = L;
    total += L;
  }
  return total;
}

void msvh_gen_residues(uint32_t seed, int64_t n, const int32_t *lens, const int64_t *off, uint8_t *res, int nthreads) {
  const int64_t nblocks = (n + 4095) / 4096;
  if (nthreads < 1) nthreads = 1;
  auto work = [&](int tid) {
    for (int64_t b = tid; b < nblocks; b += nthreads) {
      msvh_rng r;
      msvh_rng_init(&r, (uint32_t) (seed * 1000003u + (uint32_t) b + 1u), 0);
      const int64_t i1 = std::min<int64_t>(n, (b + 1) * 4096);
      for (int64_t i = b * 4096; i < i1; i++) msvh_iid(&r, lens[i], res + off[i]);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
}

}  // extern "C"
