// pressed.cpp -- reader for HMMER 3.1's pressed profiles (.h3f / .h3p, format
// 3/f), written by p7_oprofile_Write (src/impl_sse/io.c:89-180).
//
// The files hold the striped SSE vectors as raw bytes: rbv as Kp rows of
// Q16 = max(2, ceil(M/16)) 16-byte vectors (byte z of vector q is node
// q + z*Q16 + 1), rwv as Kp rows of Q8 = max(2, ceil(M/8)) vectors of eight
// int16 (word z of vector q is node q + z*Q8 + 1), twv as 7 vectors per q
// followed by the Q8 DD vectors (p7_oprofile.c:625-630, :680-715).  Fields
// follow each other without padding, in the order of io.c; the reader
// de-stripes them into the model-order tables of include/msv_gpu.h.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/msv_pressed.h"

namespace {

constexpr uint32_t kMagicF = 0xb3e6e6f3u;   // "3ffs" + 0x80808080, io.c:48
constexpr uint32_t kMagicP = 0xb3e6f0f3u;   // "3fps" + 0x80808080, io.c:49
constexpr int kOK = 0, kEOF = 3, kEMEM = 5, kEFORMAT = 7, kENOTFOUND = 6;   // easel.h:98-102

int kp_of(int alphatype) {                   // esl_alphabet.h:12-14, esl_alphabet.c:275-285
  switch (alphatype) {
    case 1: case 2: return 18;               // RNA, DNA: 4 residues + 14 other codes
    case 3: return 29;                       // amino: 20 + 9
    default: return -1;
  }
}
int nq(int M, int per) { const int q = (M - 1) / per + 1; return q < 2 ? 2 : q; }

struct Reader {
  FILE *f;
  bool ok = true;
  explicit Reader(FILE *fp) : f(fp) {}
  template <class T> T get() {
    T v{};
    if (ok && std::fread(&v, sizeof(T), 1, f) != 1) ok = false;
    return v;
  }
  void bytes(void *dst, size_t n) {
    if (ok && n && std::fread(dst, 1, n, f) != n) ok = false;
  }
  void skip(size_t n) {
    if (ok && n && std::fseek(f, (long) n, SEEK_CUR) != 0) ok = false;
  }
};

}  // namespace

struct msvh_pressed {
  FILE *f = nullptr, *p = nullptr;
  std::vector<uint8_t> rbv, raw;
  std::vector<int16_t> rwv, twv;
};

extern "C" {

int msvh_pressed_open(const char *h3f, const char *h3p, msvh_pressed **ret) {
  if (!ret || !h3f) return kENOTFOUND;
  *ret = nullptr;
  auto *ph = new (std::nothrow) msvh_pressed;
  if (!ph) return kEMEM;
  ph->f = std::fopen(h3f, "rb");
  if (h3p) ph->p = std::fopen(h3p, "rb");
  if (!ph->f || (h3p && !ph->p)) { msvh_pressed_close(ph); return kENOTFOUND; }
  *ret = ph;
  return kOK;
}

void msvh_pressed_close(msvh_pressed *ph) {
  if (!ph) return;
  if (ph->f) std::fclose(ph->f);
  if (ph->p) std::fclose(ph->p);
  delete ph;
}

int msvh_pressed_next(msvh_pressed *ph, msvg_oprofile *msv, msvg_vprofile *vit, char *name, int name_cap) {
  if (!ph || !msv) return kENOTFOUND;
  // ---- .h3f: the MSV part (io.c:97-121; read side :204-265) ----
  Reader r(ph->f);
  uint32_t magic = 0;
  if (std::fread(&magic, 4, 1, ph->f) != 1) return kEOF;
  if (magic != kMagicF) return kEFORMAT;
  const int M = r.get<int32_t>();
  const int type = r.get<int32_t>();
  const int n = r.get<int32_t>();
  if (!r.ok || M < 1 || n < 0 || n > (1 << 20)) return kEFORMAT;
  const int Kp = kp_of(type);
  if (Kp < 0) return kEFORMAT;
  std::string nm((size_t) n + 1, '\0');
  r.bytes(&nm[0], (size_t) n + 1);
  r.get<int32_t>();                                   // max_length
  const uint8_t tbm = r.get<uint8_t>(), tec = r.get<uint8_t>();
  r.get<uint8_t>();                                   // tjb_b for the length last configured
  const float scale_b = r.get<float>();
  const uint8_t base = r.get<uint8_t>(), bias = r.get<uint8_t>();
  const int Q16 = nq(M, 16), Q16x = Q16 + 17;         // p7O_EXTRA_SB, impl_sse.h:30
  r.skip((size_t) Kp * Q16x * 16);                    // sbv (the SSV copy of rbv)
  ph->raw.resize((size_t) Kp * Q16 * 16);
  r.bytes(ph->raw.data(), ph->raw.size());
  float ev[6];
  for (float &e : ev) e = r.get<float>();             // p7_NEVPARAM, hmmer.h:69
  r.skip(3 * sizeof(int64_t) + 20 * sizeof(float));   // offs (off_t), compo (p7_MAXABET)
  if (!r.ok || r.get<uint32_t>() != kMagicF) return kEFORMAT;
  ph->rbv.assign((size_t) Kp * M, 0);
  for (int x = 0; x < Kp; x++)
    for (int q = 0; q < Q16; q++)
      for (int z = 0; z < 16; z++) {
        const int k = q + z * Q16 + 1;
        if (k <= M) ph->rbv[(size_t) x * M + k - 1] = ph->raw[((size_t) x * Q16 + q) * 16 + z];
      }
  msv->M = M; msv->Kp = Kp; msv->rbv = ph->rbv.data();
  msv->tbm_b = tbm; msv->tec_b = tec; msv->base_b = base; msv->bias_b = bias; msv->scale_b = scale_b;
  msv->mmu = ev[0]; msv->mlambda = ev[1];             // p7_MMU, p7_MLAMBDA
  if (name && name_cap > 0) {
    std::strncpy(name, nm.c_str(), (size_t) name_cap - 1);
    name[name_cap - 1] = '\0';
  }
  if (!vit) return kOK;
  std::memset(vit, 0, sizeof *vit);
  if (!ph->p) return kOK;

  // ---- .h3p: the rest; only the ViterbiFilter part is kept (io.c:124-158;
  // read side :510-580) ----
  Reader s(ph->p);
  if (s.get<uint32_t>() != kMagicP) return kEFORMAT;
  const int M2 = s.get<int32_t>(), type2 = s.get<int32_t>(), n2 = s.get<int32_t>();
  if (!s.ok || M2 != M || type2 != type || n2 != n) return kEFORMAT;
  std::string nm2((size_t) n2 + 1, '\0');
  s.bytes(&nm2[0], (size_t) n2 + 1);
  if (nm2 != nm) return kEFORMAT;
  for (int t = 0; t < 2; t++) {                       // acc, desc: length, then n+1 bytes if n > 0
    const int len = s.get<int32_t>();
    if (len < 0) return kEFORMAT;
    if (len > 0) s.skip((size_t) len + 1);
  }
  s.skip(4 * ((size_t) M + 2));                       // rf, mm, cs, consensus
  const int Q8 = nq(M, 8), Q4 = nq(M, 4);
  std::vector<int16_t> tw((size_t) 8 * Q8 * 8), rw((size_t) Kp * Q8 * 8);
  s.bytes(tw.data(), tw.size() * 2);
  s.bytes(rw.data(), rw.size() * 2);
  int16_t xw[4][2];                                   // p7O_E, N, J, C x MOVE, LOOP (impl_sse.h:73-74)
  for (auto &st : xw) { st[0] = s.get<int16_t>(); st[1] = s.get<int16_t>(); }
  const float scale_w = s.get<float>();
  const int16_t base_w = s.get<int16_t>(), ddbound_w = s.get<int16_t>();
  s.get<float>();                                     // ncj_roundoff
  s.skip((size_t) 8 * Q4 * 16 + (size_t) Kp * Q4 * 16 + 4 * 2 * sizeof(float));   // tfv, rfv, xf
  s.skip(6 * sizeof(float));                          // cutoff (p7_NCUTOFFS)
  const float nj = s.get<float>();
  const int mode = s.get<int32_t>();
  s.get<int32_t>();                                   // L
  if (!s.ok || s.get<uint32_t>() != kMagicP) return kEFORMAT;
  (void) mode;
  ph->rwv.assign((size_t) Kp * M, 0);
  ph->twv.assign((size_t) 8 * M, 0);
  for (int q = 0; q < Q8; q++)
    for (int z = 0; z < 8; z++) {
      const int k = q + z * Q8 + 1;
      if (k > M) continue;
      for (int x = 0; x < Kp; x++) ph->rwv[(size_t) x * M + k - 1] = rw[((size_t) x * Q8 + q) * 8 + z];
      for (int t = 0; t < 7; t++) ph->twv[(size_t) t * M + k - 1] = tw[((size_t) q * 7 + t) * 8 + z];
      ph->twv[(size_t) 7 * M + k - 1] = tw[((size_t) 7 * Q8 + q) * 8 + z];
    }
  vit->M = M; vit->Kp = Kp; vit->rwv = ph->rwv.data(); vit->twv = ph->twv.data();
  vit->e_loop = xw[0][1]; vit->e_move = xw[0][0];
  vit->base_w = base_w; vit->ddbound_w = ddbound_w;
  vit->scale_w = scale_w; vit->nj = nj;
  vit->vmu = ev[2]; vit->vlambda = ev[3];             // p7_VMU, p7_VLAMBDA
  return kOK;
}

}  // extern "C"
