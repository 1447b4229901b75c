// hmmfile.cpp -- reader for HMMER3 ASCII profile files (declared in
// include/msv_hmmfile.h).
//
// Follows the ASCII parser of p7_hmmfile_Read (read_asc30hmm,
// src/p7_hmmfile.c:1295-1630): whitespace-separated tokens, lines whose
// first token starts with '#' and blank lines skipped (esl_fileparser.c:
// 271-301); a header of tagged lines up to "HMM"; two column-header lines;
// an optional COMPO line; node 0's insert emissions and transitions; per node
// a match line (node number, K probabilities, MAP, CONS from 3/e on, RF, MM
// from 3/f on, CS), an insert line and a transition line; then "//".
// Probabilities are stored as negative natural logs with '*' for zero; each
// is read back as expf(-atof(tok)) in float, the rounding P7_HMM holds.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <math.h>
#include <new>
#include <string>
#include <strings.h>
#include <vector>

#include "../../include/msv_hmmfile.h"
#include "../../include/msv_prep.h"

namespace {

constexpr int kOK = 0, kEOF = 3, kEMEM = 5, kENOTFOUND = 6, kEFORMAT = 7, kEINCOMPAT = 10;   // easel.h:95-106

struct Failure {
  int code;
};

}  // namespace

struct msvh_hmmfile {
  std::string buf;             // the whole file
  size_t pos = 0;              // start of the next unread line
  int lineno = 0;
  char format = 'f';           // HMMER3/<format>
  bool newly_opened = true;    // the first model's tag line was read by open
  int alphatype = 0;           // the first model's alphabet
  // the current line's tokens and the rest-of-line after each
  std::vector<std::string> tok;
  std::vector<size_t> tok_end;
  std::string line;
  size_t next_tok = 0;
  std::string err;
  // the model returned last
  std::string name, acc, desc;
  std::vector<float> mat, ins, t, compo;

  // Next line with at least one token whose first token does not start with
  // '#'; false at the end of the file.
  bool next_line() {
    while (pos < buf.size()) {
      size_t e = buf.find('\n', pos);
      if (e == std::string::npos) e = buf.size();
      line.assign(buf, pos, e - pos);
      pos = e + 1;
      ++lineno;
      tok.clear();
      tok_end.clear();
      size_t i = 0;
      while (i < line.size()) {
        while (i < line.size() && std::strchr(" \t\r", line[i])) ++i;
        if (i >= line.size()) break;
        const size_t j0 = i;
        while (i < line.size() && !std::strchr(" \t\r", line[i])) ++i;
        tok.emplace_back(line, j0, i - j0);
        tok_end.push_back(i);
      }
      if (tok.empty() || tok[0][0] == '#') continue;
      next_tok = 0;
      return true;
    }
    return false;
  }
  [[noreturn]] void fail(int code, const std::string &what) {
    err = what + " (line " + std::to_string(lineno) + ")";
    throw Failure{code};
  }
  void need_line(const char *what) {
    if (!next_line()) fail(kEOF, std::string("premature end of file: ") + what);
  }
  const std::string &token(const char *what) {
    if (next_tok >= tok.size()) fail(kEFORMAT, what);
    return tok[next_tok++];
  }
  // The rest of the line after the last token taken, leading blanks and
  // trailing blanks and line ends removed (esl_fileparser_GetRemainingLine).
  std::string rest(const char *what) {
    size_t i = next_tok ? tok_end[next_tok - 1] : 0;
    while (i < line.size() && std::strchr(" \t", line[i])) ++i;
    size_t e = line.size();
    while (e > i && std::strchr(" \t\r", line[e - 1])) --e;
    if (i >= e) fail(kEFORMAT, what);
    next_tok = tok.size();
    return line.substr(i, e - i);
  }
};

namespace {

float prob(const std::string &s) {      // p7_hmmfile.c:1535-1590
  return (s[0] == '*') ? 0.0f : expf(-1.0 * atof(s.c_str()));
}

bool yes_no(const std::string &s, bool &out) {
  if (strcasecmp(s.c_str(), "yes") == 0) { out = true; return true; }
  if (strcasecmp(s.c_str(), "no") == 0) { out = false; return true; }
  return false;
}

int encode_type(const std::string &s) {   // esl_abc_EncodeType, esl_alphabet.c:1586-1595
  if (strcasecmp(s.c_str(), "amino") == 0) return MSVH_AMINO;
  if (strcasecmp(s.c_str(), "rna") == 0) return MSVH_RNA;
  if (strcasecmp(s.c_str(), "dna") == 0) return MSVH_DNA;
  return 0;
}

int read_model(msvh_hmmfile *h, msvh_hmm *out) {
  if (!h->newly_opened) {
    if (!h->next_line()) return kEOF;                              // a normal end
    const std::string want = std::string("HMMER3/") + h->format;
    if (h->tok[0] != want) h->fail(kEFORMAT, "expected " + want + " tag, found " + h->tok[0]);
  }
  h->newly_opened = false;
  h->name.clear(); h->acc.clear(); h->desc.clear();
  int M = 0, type = 0, maxl = 0;
  bool have_name = false, seen_hmm = false, flag = false;
  float ev[6] = {0, 0, 0, 0, 0, 0};
  unsigned stats = 0;
  h->next_tok = 0;
  // header (p7_hmmfile.c:1340-1507); tags the filters do not use are parsed
  // for their format checks and dropped
  while (h->next_line()) {
    const std::string tag = h->token("no tag");
    if (tag == "NAME") { h->name = h->token("no name on NAME line"); have_name = true; }
    else if (tag == "ACC") h->acc = h->token("no accession on ACC line");
    else if (tag == "DESC") h->desc = h->rest("no description on DESC line");
    else if (tag == "LENG") {
      M = atoi(h->token("no model length on LENG line").c_str());
      if (M == 0) h->fail(kEFORMAT, "invalid model length on LENG line");
    } else if (h->format >= 'c' && tag == "MAXL") {
      maxl = atoi(h->token("no max length on MAXL line").c_str());
      if (maxl == 0) h->fail(kEFORMAT, "invalid max length on MAXL line");
    } else if (tag == "ALPH") {
      const std::string &a = h->token("no alphabet type on ALPH line");
      type = encode_type(a);
      if (!type) h->fail(kEFORMAT, "unrecognized or unsupported alphabet type " + a);
      if (h->alphatype && h->alphatype != type) h->fail(kEINCOMPAT, "alphabet type differs from the file's first model");
    } else if (tag == "RF" || tag == "MM" || tag == "CONS" || tag == "CS" || tag == "MAP") {
      const std::string &v = h->token("no yes/no on header line");
      if (!yes_no(v, flag)) h->fail(kEFORMAT, tag + " header line must say yes/no, not " + v);
    } else if (tag == "DATE") h->rest("no date on DATE line");
    else if (tag == "COM") { h->token("no command number on COM line"); h->rest("no command on COM line"); }
    else if (tag == "NSEQ") {
      if (atoi(h->token("nothing follows NSEQ").c_str()) == 0) h->fail(kEFORMAT, "invalid NSEQ");
    } else if (tag == "EFFN") {
      if (atof(h->token("nothing follows EFFN").c_str()) <= 0.0f) h->fail(kEFORMAT, "invalid EFFN");
    } else if (tag == "CKSUM") h->token("nothing follows CKSUM");
    else if (tag == "STATS") {
      if (h->format >= 'b') {                                      // STATS LOCAL MSV|VITERBI|FORWARD mu lambda
        const std::string t1 = h->token("too few fields on STATS line"), t2 = h->token("too few fields on STATS line");
        const std::string t3 = h->token("too few fields on STATS line"), t4 = h->token("too few fields on STATS line");
        if (strcasecmp(t1.c_str(), "LOCAL") != 0) h->fail(kEFORMAT, "STATS: " + t1 + " unrecognized as field 2");
        int which;
        if (strcasecmp(t2.c_str(), "MSV") == 0) which = 0;
        else if (strcasecmp(t2.c_str(), "VITERBI") == 0) which = 1;
        else if (strcasecmp(t2.c_str(), "FORWARD") == 0) which = 2;
        else h->fail(kEFORMAT, "STATS: " + t2 + " unrecognized as field 3");
        ev[2 * which] = atof(t3.c_str());
        ev[2 * which + 1] = atof(t4.c_str());
        stats |= 1u << which;
      } else {                                                     // 3/a: STATS LOCAL VLAMBDA|VMU|FTAU value
        const std::string t1 = h->token("too few fields on STATS line"), t2 = h->token("too few fields on STATS line");
        const std::string t3 = h->token("too few fields on STATS line");
        if (strcasecmp(t1.c_str(), "LOCAL") != 0) h->fail(kEFORMAT, "STATS: " + t1 + " unrecognized as field 2");
        const float v = atof(t3.c_str());
        if (strcasecmp(t2.c_str(), "VLAMBDA") == 0) { ev[MSVH_MLAMBDA] = ev[MSVH_VLAMBDA] = ev[MSVH_FLAMBDA] = v; stats |= 1; }
        else if (strcasecmp(t2.c_str(), "VMU") == 0) { ev[MSVH_MMU] = ev[MSVH_VMU] = v; stats |= 2; }
        else if (strcasecmp(t2.c_str(), "FTAU") == 0) { ev[MSVH_FTAU] = v; stats |= 4; }
        else h->fail(kEFORMAT, "STATS: " + t2 + " unrecognized as field 3");
      }
    } else if (tag == "GA" || tag == "TC" || tag == "NC") {
      h->token("too few fields on cutoff line");
      if (type != MSVH_DNA && type != MSVH_RNA) h->token("too few fields on cutoff line");
    } else if (tag == "HMM") { seen_hmm = true; break; }
  }
  if (!seen_hmm) h->fail(kEOF, "premature end of file in a model's header");
  if (stats != 0 && stats != 7) h->fail(kEFORMAT, "missing one or more STATS parameter lines");
  if (M <= 0) h->fail(kEFORMAT, "no LENG found for HMM (or LENG <= 0)");
  if (!type) h->fail(kEFORMAT, "no ALPH found for HMM");
  const int K = (type == MSVH_AMINO) ? 20 : 4;

  h->need_line("before the main model section");                  // the m->m m->i ... line
  h->need_line("before the main model section");
  h->mat.assign((size_t) (M + 1) * K, 0.0f);
  h->ins.assign((size_t) (M + 1) * K, 0.0f);
  h->t.assign((size_t) (M + 1) * 7, 0.0f);
  h->compo.clear();
  std::string t0 = h->token("premature end of data before main model section");
  if (t0 == "COMPO") {                                             // :1520-1529
    h->compo.resize((size_t) K);
    for (int x = 0; x < K; x++) h->compo[(size_t) x] = prob(h->token("too few fields on COMPO line"));
    h->need_line("after COMPO line");
    t0 = h->token("premature end of data after COMPO line");
  }
  h->ins[0] = prob(t0);                                            // node 0, :1531-1542
  for (int x = 1; x < K; x++) h->ins[(size_t) x] = prob(h->token("too few fields on insert line, node 0"));
  h->need_line("no node 0 transition line");
  for (int x = 0; x < 7; x++) h->t[(size_t) x] = prob(h->token("too few fields on begin (0) transition line"));
  for (int k = 1; k <= M; k++) {                                   // :1545-1597
    const std::string ks = std::to_string(k);
    h->need_line(("in the main model section, at node " + ks).c_str());
    if (atoi(h->token("premature end of data in main model section").c_str()) != k)
      h->fail(kEFORMAT, "expected match line to start with " + ks);
    for (int x = 0; x < K; x++)
      h->mat[(size_t) k * K + x] = prob(h->token(("too few probability fields on match line, node " + ks).c_str()));
    h->token(("missing MAP field on match line for node " + ks).c_str());
    if (h->format >= 'e') h->token(("missing CONS field on match line for node " + ks).c_str());
    h->token(("missing RF field on match line for node " + ks).c_str());
    if (h->format >= 'f') h->token(("missing MM field on match line for node " + ks).c_str());
    h->token(("missing CS field on match line for node " + ks).c_str());
    h->need_line(("no insert emission line, node " + ks).c_str());
    for (int x = 0; x < K; x++)
      h->ins[(size_t) k * K + x] = prob(h->token(("too few probability fields on insert line, node " + ks).c_str()));
    h->need_line(("no transition line, node " + ks).c_str());
    for (int x = 0; x < 7; x++)
      h->t[(size_t) k * 7 + x] = prob(h->token(("too few probability fields on transition line, node " + ks).c_str()));
  }
  h->need_line("missing //?");
  if (h->token("missing //?") != "//") h->fail(kEFORMAT, "expected closing //; found " + h->tok[0]);
  if (!have_name) h->fail(kEFORMAT, "no NAME found for HMM");
  if (!h->alphatype) h->alphatype = type;

  out->M = M;
  out->alphatype = type;
  out->K = K;
  out->Kp = (type == MSVH_AMINO) ? 29 : 18;
  out->mat = h->mat.data();
  out->ins = h->ins.data();
  out->t = h->t.data();
  out->compo = h->compo.empty() ? nullptr : h->compo.data();
  for (int i = 0; i < 6; i++) out->evparam[i] = (stats == 7) ? ev[i] : 0.0f;
  out->has_stats = (stats == 7);
  out->max_length = maxl;
  out->name = h->name.c_str();
  out->acc = h->acc.c_str();
  out->desc = h->desc.c_str();
  return kOK;
}

}  // namespace

extern "C" {

int msvh_hmmfile_open(const char *path, msvh_hmmfile **ret) {
  if (!ret) return kENOTFOUND;
  *ret = nullptr;
  if (!path) return kENOTFOUND;
  FILE *f = std::fopen(path, "rb");
  if (!f) return kENOTFOUND;
  auto *h = new (std::nothrow) msvh_hmmfile;
  if (!h) { std::fclose(f); return kEMEM; }
  char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof chunk, f)) > 0) h->buf.append(chunk, n);
  std::fclose(f);
  if (!h->next_line() || h->tok[0].size() != 8 || h->tok[0].compare(0, 7, "HMMER3/") != 0 || h->tok[0][7] < 'a' ||
      h->tok[0][7] > 'f') {
    delete h;
    return kEFORMAT;                                             // p7_hmmfile.c:451-458
  }
  h->format = h->tok[0][7];
  *ret = h;
  return kOK;
}

int msvh_hmmfile_next(msvh_hmmfile *h, msvh_hmm *out) {
  if (!h || !out) return kENOTFOUND;
  h->err.clear();
  try {
    return read_model(h, out);
  } catch (const Failure &e) {
    return e.code;
  } catch (const std::bad_alloc &) {
    h->err = "allocation failed";
    return kEMEM;
  }
}

const char *msvh_hmmfile_error(const msvh_hmmfile *h) { return h ? h->err.c_str() : ""; }

void msvh_hmmfile_close(msvh_hmmfile *h) { delete h; }

}  // extern "C"
