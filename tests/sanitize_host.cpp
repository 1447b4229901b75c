// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
// HMMER3 ASCII reader, the pressed-file reader, the profile preparation
// (sampling, MSV and ViterbiFilter conversions, calibration fits, target
// generators) and the CPU oracle, exercised on the golden model files, on
// truncated and corrupted copies of them and on sampled models and targets.
// Built and run by tests/test_sanitized_host.py with
// g++ -fsanitize=address,undefined; a finding aborts the run, a failed
// consistency check makes it exit 1.
//
//   sanitize_host hmm <file.hmm>...        read, convert, score every model
//   sanitize_host pressed <h3f> <h3p>      read every pressed model and score it
//   sanitize_host mangle <file.hmm> <tmp>  truncated / corrupted copies of a file
//   sanitize_host prep                     sampled models, conversions, fits, generators
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "msv_gpu.h"
#include "msv_hmmfile.h"
#include "msv_prep.h"
#include "msv_pressed.h"
extern "C" {
#include "msv_oracle.h"
#include "vit_oracle.h"
}

static int g_fails = 0;
#define CHECK(cond, ...)                                     \
  do {                                                       \
    if (!(cond)) {                                           \
      std::fprintf(stderr, "check failed: " __VA_ARGS__);    \
      std::fputc('\n', stderr);                              \
      ++g_fails;                                             \
    }                                                        \
  } while (0)

// Targets of several lengths (the empty one, a 16-residue block's edges,
// a long one) with residue codes below K and a few degenerate codes.
static std::vector<std::vector<uint8_t>> make_targets(int K, int Kp, uint32_t seed) {
  msvh_rng r;
  msvh_rng_init(&r, seed, 1);
  std::vector<std::vector<uint8_t>> out;
  for (int L : {0, 1, 2, 15, 16, 17, 31, 300, 3001}) {
    std::vector<uint8_t> x((size_t) L);
    for (int i = 0; i < L; ++i) {
      const double u = msvh_random(&r);
      x[(size_t) i] = (uint8_t) (u < 0.02 && Kp > K + 3 ? K + 1 + (int) (u * 50.0) % (Kp - K - 3) : (int) (u * K) % K);
    }
    out.push_back(std::move(x));
  }
  return out;
}

// MSV and ViterbiFilter conversions of one model, then the oracle on the
// targets: each target alone equals the batch call, statuses are valid.
static void score_model(int alphatype, int M, int K, int Kp, const float *mat, const float *t, uint32_t seed) {
  std::vector<uint8_t> rbv((size_t) Kp * M);
  int sc[5] = {0};
  float scale = 0.f;
  int rc = msvh_profile_convert_abc(alphatype, M, mat, 100, rbv.data(), sc, &scale, nullptr);
  CHECK(rc == 0, "msvh_profile_convert_abc rc %d (M %d)", rc, M);
  if (rc != 0) return;
  const double lam = msvh_lambda_abc(alphatype, M, mat);
  CHECK(std::isfinite(lam) && lam > 0.0, "lambda %g", lam);
  orc_profile p{M, Kp, rbv.data(), sc[0], sc[1], sc[3], sc[4], scale};

  std::vector<int16_t> rwv((size_t) Kp * M), twv((size_t) 8 * M);
  int vs[4] = {0};
  float sw = 0.f, nj = 0.f;
  rc = msvh_vit_convert_abc(alphatype, M, mat, t, rwv.data(), twv.data(), vs, &sw, &nj);
  CHECK(rc == 0, "msvh_vit_convert_abc rc %d (M %d)", rc, M);
  orc_vprofile vp{M, Kp, rwv.data(), twv.data(), vs[0], vs[1], vs[2], vs[3], sw, nj};
  // the ViterbiFilter planner's lazy-F estimate over the same fields
  const double lr = msvh_vit_lazyf_rate(rwv.data(), twv.data(), M, Kp, vs[0], vs[2], vs[3],
                                        msvh_vit_move(sw, nj, 350), 64);
  CHECK(lr >= 0.0 && lr <= 1.0, "lazy-F rate %g (M %d)", lr, M);

  const auto tg = make_targets(K, Kp, seed);
  std::vector<uint8_t> res;
  std::vector<int64_t> off;
  std::vector<int32_t> len;
  std::vector<float> usc1, vsc1;
  std::vector<int32_t> st1, vst1;
  for (const auto &x : tg) {
    off.push_back((int64_t) res.size());
    len.push_back((int32_t) x.size());
    res.insert(res.end(), x.begin(), x.end());
    float u = 0.f, v = 0.f;
    const int s = orc_msv(&p, x.data(), (int) x.size(), &u);
    CHECK(s == ORC_eslOK || s == ORC_eslERANGE, "orc_msv status %d", s);
    usc1.push_back(u);
    st1.push_back(s);
    const int vstat = (rc == 0) ? orc_vit(&vp, x.data(), (int) x.size(), &v) : 0;
    CHECK(vstat == ORC_eslOK || vstat == ORC_eslERANGE, "orc_vit status %d", vstat);
    vsc1.push_back(v);
    vst1.push_back(vstat);
    const float nul = msvh_nullsc((int) x.size());   // L = 0: 0 * log(0), as the reference (never scored)
    CHECK(x.empty() || std::isfinite(nul), "nullsc(%zu)", x.size());
    (void) orc_pass(u, (int) x.size(), -8.0f, (float) lam, 0.02);
    (void) msvh_tjb((int) x.size());
  }
  res.resize(res.size() + 16, 0);
  const int n = (int) tg.size();
  std::vector<float> usc2((size_t) n), vsc2((size_t) n);
  std::vector<int32_t> st2((size_t) n), vst2((size_t) n);
  orc_msv_batch(&p, res.data(), off.data(), len.data(), n, usc2.data(), st2.data(), 2);
  if (rc == 0) orc_vit_batch(&vp, res.data(), off.data(), len.data(), n, vsc2.data(), vst2.data(), 2);
  for (int i = 0; i < n; ++i) {
    CHECK(st1[(size_t) i] == st2[(size_t) i] && std::memcmp(&usc1[(size_t) i], &usc2[(size_t) i], 4) == 0,
          "MSV target %d: one-by-one and batch differ", i);
    if (rc == 0)
      CHECK(vst1[(size_t) i] == vst2[(size_t) i] && std::memcmp(&vsc1[(size_t) i], &vsc2[(size_t) i], 4) == 0,
            "ViterbiFilter target %d: one-by-one and batch differ", i);
  }
}

static int run_hmm(int argc, char **argv) {
  int models = 0;
  for (int a = 0; a < argc; ++a) {
    msvh_hmmfile *f = nullptr;
    int rc = msvh_hmmfile_open(argv[a], &f);
    CHECK(rc == 0, "open %s: rc %d", argv[a], rc);
    if (rc != 0) continue;
    msvh_hmm h;
    while ((rc = msvh_hmmfile_next(f, &h)) == 0) {
      CHECK(h.M >= 1 && h.K >= 4 && h.Kp > h.K && h.name && h.name[0], "model header of %s", argv[a]);
      score_model(h.alphatype, h.M, h.K, h.Kp, h.mat, h.t, 1000u + (uint32_t) models);
      ++models;
    }
    CHECK(rc == 3, "%s ended with rc %d (%s)", argv[a], rc, msvh_hmmfile_error(f));
    msvh_hmmfile_close(f);
  }
  std::printf("hmm: %d models read, converted and scored\n", models);
  return 0;
}

static int run_pressed(const char *h3f, const char *h3p) {
  msvh_pressed *ph = nullptr;
  int rc = msvh_pressed_open(h3f, h3p, &ph);
  CHECK(rc == 0, "pressed open rc %d", rc);
  if (rc != 0) return 0;
  msvg_oprofile msv;
  msvg_vprofile vit;
  char name[64];
  int models = 0;
  while ((rc = msvh_pressed_next(ph, &msv, &vit, name, (int) sizeof name)) == 0) {
    CHECK(msv.M >= 1 && msv.Kp >= 18, "pressed model %d", models);
    orc_profile p{msv.M, msv.Kp, msv.rbv, msv.tbm_b, msv.tec_b, msv.base_b, msv.bias_b, msv.scale_b};
    orc_vprofile vp{vit.M, vit.Kp, vit.rwv, vit.twv, vit.e_loop, vit.e_move, vit.base_w, vit.ddbound_w,
                    vit.scale_w, vit.nj};
    const int K = msv.Kp == 29 ? 20 : 4;
    for (const auto &x : make_targets(K, msv.Kp, 77u + (uint32_t) models)) {
      float u, v;
      const int s = orc_msv(&p, x.data(), (int) x.size(), &u);
      CHECK(s == ORC_eslOK || s == ORC_eslERANGE, "pressed orc_msv status %d", s);
      if (vit.M == msv.M) {
        const int t = orc_vit(&vp, x.data(), (int) x.size(), &v);
        CHECK(t == ORC_eslOK || t == ORC_eslERANGE, "pressed orc_vit status %d", t);
      }
    }
    ++models;
  }
  CHECK(rc == 3, "pressed file ended with rc %d", rc);
  msvh_pressed_close(ph);
  rc = msvh_pressed_open("/nonexistent/x.h3f", nullptr, &ph);
  CHECK(rc != 0, "a missing pressed file opened");
  std::printf("pressed: %d models read and scored\n", models);
  return 0;
}

// Every prefix length on a coarse grid, then single corrupted bytes: the
// reader returns a status for each copy (or reads what is intact) and never
// touches memory it does not own.
static int run_mangle(const char *path, const char *tmp) {
  std::FILE *fp = std::fopen(path, "rb");
  CHECK(fp != nullptr, "open %s", path);
  if (!fp) return 0;
  std::string data;
  char buf[65536];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof buf, fp)) > 0) data.append(buf, k);
  std::fclose(fp);
  auto read_all = [&](const std::string &bytes) {
    std::FILE *o = std::fopen(tmp, "wb");
    std::fwrite(bytes.data(), 1, bytes.size(), o);
    std::fclose(o);
    msvh_hmmfile *f = nullptr;
    if (msvh_hmmfile_open(tmp, &f) != 0) return;
    msvh_hmm h;
    int rc, n = 0;
    while ((rc = msvh_hmmfile_next(f, &h)) == 0 && n < 64) {
      CHECK(h.M >= 1, "a model with M %d", h.M);
      ++n;
    }
    CHECK(rc == 0 || rc == 3 || rc == 7 || rc == 10 || rc == 11, "status %d", rc);
    msvh_hmmfile_close(f);
  };
  const size_t step = data.size() / 97 + 1;
  int copies = 0;
  for (size_t cut = 0; cut < data.size(); cut += step, ++copies) read_all(data.substr(0, cut));
  msvh_rng r;
  msvh_rng_init(&r, 4242u, 1);
  const char junk[] = {'\n', ' ', '*', '-', '9', 'x', '\0', '\t', '.', 'e'};
  for (int i = 0; i < 120; ++i, ++copies) {
    std::string c = data;
    const size_t at = (size_t) (msvh_random(&r) * (double) c.size()) % c.size();
    c[at] = junk[(size_t) (msvh_random(&r) * 10.0) % 10];
    read_all(c);
  }
  std::printf("mangle: %d damaged copies of %s read\n", copies, path);
  return 0;
}

static int run_prep() {
  msvh_rng r;
  msvh_rng_init(&r, 7u, 0);
  int models = 0;
  for (int M : {1, 2, 7, 64, 400, 2500}) {
    std::vector<float> mat((size_t) (M + 1) * 20), ins((size_t) (M + 1) * 20), t((size_t) (M + 1) * 7);
    int rc = msvh_hmm_sample(&r, M, mat.data(), ins.data(), t.data());
    CHECK(rc == 0, "msvh_hmm_sample rc %d", rc);
    score_model(MSVH_AMINO, M, 20, 29, mat.data(), t.data(), 10u + (uint32_t) M);
    std::vector<uint8_t> em(40000);
    const int L = msvh_core_emit(&r, M, mat.data(), ins.data(), t.data(), em.data(), (int) em.size());
    CHECK(L >= 0, "core_emit cap");
    const int Ls = msvh_core_emit(&r, M, mat.data(), ins.data(), t.data(), em.data(), 1);
    CHECK(Ls >= -1 && Ls <= 1, "core_emit into one byte: %d", Ls);
    ++models;
  }
  // calibration targets and the two fits on made-up scores (ERANGE included)
  const int N = 200, L = 100;
  std::vector<uint8_t> ct((size_t) N * L);
  msvh_calibration_targets(N, L, ct.data());
  msvh_vit_calibration_targets(N, L, ct.data());
  std::vector<float> sc((size_t) N);
  std::vector<int32_t> st((size_t) N);
  for (int i = 0; i < N; ++i) {
    sc[(size_t) i] = (float) (msvh_random(&r) * 10.0 - 5.0);
    st[(size_t) i] = (i % 37 == 0) ? 16 : 0;
  }
  const double mu = msvh_fit_mu(sc.data(), st.data(), N, L, 190, 3.0f, 0.69);
  const double vmu = msvh_fit_vmu(sc.data(), st.data(), N, L, 12000, 500.0f, 0.69);
  CHECK(std::isfinite(mu) && std::isfinite(vmu), "fits %g %g", mu, vmu);
  // generators, with and without a Pareto tail
  for (double tail : {0.0, 0.05}) {
    const int64_t n = 3000;
    std::vector<int32_t> lens((size_t) n);
    const int64_t total = msvh_gen_lengths(11u, n, 300.0, 0.6, 20, 35000, tail, 1000.0, 1.5, 35000, lens.data());
    CHECK(total > 0, "gen_lengths total %lld", (long long) total);
    std::vector<int64_t> off((size_t) n);
    int64_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
      CHECK(lens[(size_t) i] >= 20 && lens[(size_t) i] <= 35000, "length %d", lens[(size_t) i]);
      off[(size_t) i] = acc;
      acc += lens[(size_t) i];
    }
    CHECK(acc == total, "lengths sum");
    std::vector<uint8_t> resid((size_t) acc);
    msvh_gen_residues(12u, n, lens.data(), off.data(), resid.data(), 3);
    for (uint8_t x : resid) CHECK(x < 20, "residue code %d", x);
  }
  for (int Lm : {0, 1, 100, 100000}) {
    (void) msvh_vit_move(500.0f, 1.0f, Lm);
    (void) msvh_vit_score(-32768, 0, 12000, 500.0f);
    (void) msvh_vit_score(15000, -3, 12000, 500.0f);
  }
  std::printf("prep: %d sampled models, fits, generators\n", models);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: sanitize_host hmm|pressed|mangle|prep ...\n");
    return 2;
  }
  const std::string cmd = argv[1];
  if (cmd == "hmm") run_hmm(argc - 2, argv + 2);
  else if (cmd == "pressed" && argc == 4) run_pressed(argv[2], argv[3]);
  else if (cmd == "mangle" && argc == 4) run_mangle(argv[2], argv[3]);
  else if (cmd == "prep") run_prep();
  else {
    std::fprintf(stderr, "bad arguments\n");
    return 2;
  }
  if (g_fails) std::fprintf(stderr, "%d checks failed\n", g_fails);
  return g_fails ? 1 : 0;
}
