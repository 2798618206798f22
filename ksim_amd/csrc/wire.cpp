// wire.cpp — serialise packed programs and validate them on the way back in.
#include "wire.hpp"

#include <cstring>
#include <stdexcept>

namespace ksim {

namespace {

const char kMagic[8] = {'K', 'S', 'I', 'M', 'P', 'R', 'G', '\0'};

[[noreturn]] void bad(const std::string &what) {
  throw std::invalid_argument("program blob rejected: " + what);
}

template <class T>
T rd(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

template <class T>
void wr(uint8_t *p, T v) {
  std::memcpy(p, &v, sizeof(T));
}

}  // namespace

uint64_t fnv1a64(const void *data, size_t n) {
  const uint8_t *p = (const uint8_t *)data;
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

std::vector<uint8_t> blob_from_words(const std::vector<uint32_t> &words, const double *stats8) {
  if (words.size() < PH_DWORDS) throw std::invalid_argument("program words too short");
  std::vector<uint8_t> b(WIRE_HEADER + words.size() * 4, 0);
  std::memcpy(b.data(), kMagic, 8);
  wr<uint32_t>(&b[8], WIRE_VERSION);
  wr<uint32_t>(&b[12], (uint32_t)WIRE_HEADER);
  wr<uint64_t>(&b[16], (uint64_t)words.size());
  wr<uint64_t>(&b[24], fnv1a64(words.data(), words.size() * 4));
  for (int k = 0; k < 8; ++k) wr<double>(&b[32 + 8 * k], stats8 ? stats8[k] : 0.0);
  const uint32_t hdr[5] = {words[PH_NCOMP], words[PH_NUMF], words[PH_NUMOUT], words[PH_TOTOUT],
                           words[PH_WMAX]};
  for (int k = 0; k < 5; ++k) wr<uint32_t>(&b[96 + 4 * k], hdr[k]);
  std::memcpy(&b[WIRE_HEADER], words.data(), words.size() * 4);
  return b;
}

void validate_words(const uint32_t *w, size_t n) {
  auto need = [&](uint64_t end, const char *what) {
    if (end > n) bad(std::string(what) + " runs past the program words");
  };
  need(PH_DWORDS, "program header");
  const uint32_t nc = w[PH_NCOMP], num_f = w[PH_NUMF], num_out = w[PH_NUMOUT];
  const uint32_t tot = w[PH_TOTOUT], wmax = w[PH_WMAX];
  if (wmax < 1 || wmax > (uint32_t)LANE_WORDS_MAX) bad("lane words outside 1..LANE_WORDS_MAX");
  need((uint64_t)PH_DWORDS + nc, "component offsets");
  uint64_t outs = 0;
  bool all_narrow = true, any_approx = false, any_table = false;
  uint32_t table_w32 = 0;  // the widest graph with parity tables
  for (uint32_t c = 0; c < nc; ++c) {
    const uint64_t co = w[PH_DWORDS + c];
    need(co + CH_DWORDS, "component header");
    const uint32_t *ch = w + co;
    const uint32_t ng = ch[CH_NGRAPHS], F = ch[CH_F], nout = ch[CH_NOUT], cw = ch[CH_WMAX];
    if (ng < 1 || nout != ng - 1) bad("graph count / output count mismatch");
    if (cw < 1 || cw > wmax) bad("component lane words");
    if ((uint64_t)F + nout > 32ull * cw) bad("parameters exceed the lane words");
    if ((uint64_t)ch[CH_COLBASE] + nout > tot) bad("output column base");
    outs += nout;
    need((uint64_t)ch[CH_FSEL_OFF] + F, "f_selection");
    for (uint32_t j = 0; j < F; ++j)
      if (w[ch[CH_FSEL_OFF] + j] >= num_f) bad("f_selection entry beyond num_f");
    if (ch[CH_FSEL0] != 0xffffffffu) {
      if ((uint64_t)ch[CH_FSEL0] + F > num_f) bad("contiguous f_selection beyond num_f");
      for (uint32_t j = 0; j < F; ++j)
        if (w[ch[CH_FSEL_OFF] + j] != ch[CH_FSEL0] + j) bad("f_selection is not contiguous");
    }
    need((uint64_t)ch[CH_DEST_OFF] + nout, "output columns");
    for (uint32_t i = 0; i < nout; ++i)
      if (w[ch[CH_DEST_OFF] + i] >= num_out) bad("output column beyond num_outputs");
    if (ch[CH_DEST0] != 0xffffffffu) {
      if ((uint64_t)ch[CH_DEST0] + nout > num_out) bad("contiguous output columns beyond num_outputs");
      for (uint32_t i = 0; i < nout; ++i)
        if (w[ch[CH_DEST_OFF] + i] != ch[CH_DEST0] + i) bad("output columns are not contiguous");
    }
    need((uint64_t)ch[CH_GRAPH_OFF] + (uint64_t)ng * GH_DWORDS, "graph headers");
    for (uint32_t g = 0; g < ng; ++g) {
      const uint32_t *gh = w + ch[CH_GRAPH_OFF] + (size_t)g * GH_DWORDS;
      const uint32_t W32 = gh[GH_W32];
      if (W32 < 1 || W32 > cw) bad("graph mask words");
      if (W32 != std::max<uint32_t>(1u, (F + g + 31) / 32)) bad("graph width is not ceil((F + i) / 32)");
      const uint32_t nt = gh[GH_NTERMS];
      if (gh[GH_FLAGS] & ~(GF_MONO | GF_NARROW | GF_APPROX | GF_STREAM | GF_TABLE))
        bad("graph flags");
      if ((gh[GH_FLAGS] & GF_MONO) && nt != 1) bad("single-term flag on a graph of several terms");
      if (!(gh[GH_FLAGS] & (GF_MONO | GF_NARROW))) all_narrow = false;
      if (gh[GH_FLAGS] & GF_APPROX) any_approx = true;
      const bool table = gh[GH_FLAGS] & GF_TABLE;
      if (table) {
        any_table = true;
        table_w32 = std::max(table_w32, W32);
        if (W32 > (uint32_t)PM_TAB_W32_MAX) bad("parity tables on a graph wider than their limit");
        if (gh[GH_FLAGS] & (GF_MONO | GF_APPROX)) bad("parity tables on a single-term or approximate graph");
      }
      if (gh[GH_FLAGS] & GF_APPROX) {
        if (gh[GH_FLAGS] & (GF_MONO | GF_NARROW)) bad("approximate graph with an exact-sum flag");
        if (gh[GH_PART_T1] != nt) bad("approximate graph split into term parts");
        need((uint64_t)gh[GH_APX_OFF] + 2ull * nt, "approximate factors");
      }
      const uint32_t pt[GH_PARTS + 1] = {0, gh[GH_PART_T1], gh[GH_PART_T2], gh[GH_PART_T3], nt};
      for (int p = 0; p < GH_PARTS; ++p)
        if (pt[p] > pt[p + 1]) bad("term parts out of order");
      uint64_t to = gh[GH_TERM_OFF];
      for (uint32_t t = 0; t <= nt; ++t) {
        for (int p = 1; p < GH_PARTS; ++p)
          if (pt[p] == t && gh[GH_PART_O1 + p - 1] != to) bad("term part offset");
        if (t == nt) break;
        need(to + TH_DWORDS, "term header");
        const uint32_t *th = w + to;
        if (th[TH_FLAGS] & ~(TF_RESIDUAL | TF_PI)) bad("term flags");
        if ((gh[GH_FLAGS] & GF_MONO) && th[TH_FLAGS]) bad("single-term graph with pi or residual");
        uint64_t sz = (uint64_t)th[TH_NPROJ] * (W32 + 1) + (uint64_t)th[TH_NPPX] * (2 * W32 + 2);
        uint64_t rp = to + TH_DWORDS + sz;
        for (uint32_t k = 0; k < th[TH_NCHUNK]; ++k) {
          need(rp + CHUNK_HDR, "record chunk");
          const uint32_t nw = w[rp + 3];
          const uint64_t nl = w[rp], np = w[rp + 1], nm = nw & PM_COUNT_MASK;
          const uint64_t pad = nw >> PM_PAD_SHIFT;
          // (in parity-group graphs nm counts groups, not sign records)
          if (nl + np + (table ? 0 : nm) > (uint64_t)CHUNK_MAX)
            bad("record chunk longer than its limit");
          if (pad && !(table && nm)) bad("padding in a chunk without parity tables");
          if (table && nm && (pad > 3 || (rp + CHUNK_HDR + pad + PM_TAB_HDR) % 4 != 0))
            bad("parity tables not on a 16-byte boundary");
          const uint64_t pm = table ? nm * pm_tab_group_dwords(W32) : pm_dwords(nm, W32);
          const uint64_t cs = CHUNK_HDR + pad + pm + nl * (W32 + 1) + np * (2 * W32 + 3);
          sz += cs;
          rp += cs;
        }
        if (sz != th[TH_SIZE]) bad("term size does not match its records");
        need(to + TH_DWORDS + sz, "term records");
        to += TH_DWORDS + sz;
      }
      // the record read-ahead reaches STREAM_AHEAD dwords past the last record
      if (gh[GH_FLAGS] & GF_STREAM) need(to + STREAM_AHEAD, "record read-ahead padding");
    }
  }
  if (outs != tot) bad("total outputs");
  // parity tables are read by the register-word kernels that hold a buffer
  // for them: graphs of up to PM_TAB_W32_MAX words in programs with 128-bit
  // sums, of up to PM_TAB_W32_MAX_NARROW words in programs without
  if (any_table && (any_approx || wmax > (uint32_t)REG_WORDS_MAX ||
                    (all_narrow && (wmax > (uint32_t)PM_TAB_NARROW_WL ||
                                    table_w32 > (uint32_t)PM_TAB_W32_MAX_NARROW))))
    bad("parity tables in a program whose kernels cannot read them");
}

std::vector<uint32_t> words_from_blob(const uint8_t *blob, size_t n, double *stats8) {
  if (!blob) bad("NULL blob");
  if (n < WIRE_HEADER) bad("shorter than its header");
  if (std::memcmp(blob, kMagic, 8) != 0) bad("bad magic");
  const uint32_t version = rd<uint32_t>(blob + 8);
  if (version != WIRE_VERSION && version != 5 && version != 4 && version != 3)
    bad("unsupported version");
  if (rd<uint32_t>(blob + 12) != WIRE_HEADER) bad("header size");
  const uint64_t nw = rd<uint64_t>(blob + 16);
  if (nw < PH_DWORDS || nw > (n - WIRE_HEADER) / 4 || WIRE_HEADER + nw * 4 != n)
    bad("word count does not match the blob size");
  std::vector<uint32_t> words((size_t)nw);
  std::memcpy(words.data(), blob + WIRE_HEADER, (size_t)nw * 4);
  if (fnv1a64(words.data(), (size_t)nw * 4) != rd<uint64_t>(blob + 24)) bad("checksum mismatch");
  const uint32_t hdr[5] = {words[PH_NCOMP], words[PH_NUMF], words[PH_NUMOUT], words[PH_TOTOUT],
                           words[PH_WMAX]};
  for (int k = 0; k < 5; ++k)
    if (rd<uint32_t>(blob + 96 + 4 * k) != hdr[k]) bad("header fields disagree with the words");
  validate_words(words.data(), words.size());
  if (stats8)
    for (int k = 0; k < 8; ++k) stats8[k] = rd<double>(blob + 32 + 8 * k);
  return words;
}

}  // namespace ksim
