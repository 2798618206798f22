// layout.hpp — device program format shared by the host packer and the
// gfx950 kernels.
//
// Every factor of a term (the node-phase, half-pi, pi-product and phase-pair
// families of src/cpp/ksim/zx_eval.hpp:8-14) is, for each parity outcome,
// either 0 or a product
//
//     w^s (1+w)^v (1+sqrt2)^n pi^a pibar^b,   pi = 1 + i sqrt2, pibar = 1 - i sqrt2,
//
// so a term's product is fixed by five small exponents.  The kernel adds
// per-factor exponent deltas into one packed 32-bit word instead of
// multiplying four-coefficient Z[w] values, and builds the exact Z[w] value
// once per term from a table.
//
// Packed exponent word (a "chunk" of at most CHUNK_MAX factor records):
//   bits  0.. 8  V            (v <= 6 per record)
//   bits  9..16  N + NBIAS    (n in [-3, 0] per record)
//   bits 17..22  A
//   bits 23..28  B
//   bits 29..31  S mod 8      (wraps by itself)
#pragma once
#include <cstdint>

namespace ksim {

constexpr int PK_OV = 0, PK_WV = 9;
constexpr int PK_ON = 9, PK_WN = 8;
constexpr int PK_OA = 17, PK_WA = 6;
constexpr int PK_OB = 23, PK_WB = 6;
constexpr int PK_OS = 29;
constexpr int NBIAS = 3;
constexpr int CHUNK_MAX = 63;

// Materialisation table: w^s (1+w)^r (1+sqrt2)^n for |n| <= TAB_W, r < 4,
// s < 8, coefficients in the basis (1, w, i, conj w), each written as
// TAB_DIGITS balanced base-2^31 digits (every digit below 2^30 in magnitude)
// held in TAB_DIGITS planes of TAB_N int32 x 4 entries.  Rows |n| <= TAB_R
// are a single digit (planes 1.. are zero there).
constexpr int TAB_R = 22;
constexpr int TAB_W = 64;
constexpr int TAB_DIGITS = 3;
constexpr int TAB_DIGIT_BITS = 31;
constexpr int TAB_N = (2 * TAB_W + 1) * 4 * 8;  // entries per digit plane
inline int tab_index(int n, int r, int s) { return ((n + TAB_W) * 4 + r) * 8 + s; }
// An all-zero table entry: plane 1 on the single-digit row n = 0.  Lanes whose
// term vanished read it, so their addend is zero without a select per
// coefficient.
constexpr int TAB_ZERO = TAB_N + ((0 + TAB_W) * 4 + 0) * 8 + 0;
// float64 magnitudes |w^s (1+w)^r (1+sqrt2)^n| of the table entries (by the
// sampler's magnitude sequence), same indexing; rows |n| <= MAG_LDS_R in LDS
constexpr int MAG_LDS_R = 12;
constexpr int MAG_LDS_N = (2 * MAG_LDS_R + 1) * 4 * 8;

// pi^a pibar^b for a, b < PI_MAX (int32 coefficients, < 2^24)
constexpr int PI_MAX = 16;

// 32-bit parameter words a lane holds for one shot: up to REG_WORDS_MAX in
// registers (kernels specialised on 1, 2, 4 or 8 words), wider components
// (at most LANE_WORDS_MAX words, F + n_out <= 4096 parameters) in LDS
constexpr int REG_WORDS_MAX = 8;
constexpr int LANE_WORDS_MAX = 128;

// term header (dwords)
enum TermHdr : int {
  TH_NPROJ = 0, TH_NPPX, TH_NCHUNK, TH_FLAGS,
  TH_S, TH_V, TH_N, TH_A, TH_B, TH_E,
  TH_R0, TH_R1, TH_R2, TH_R3,
  TH_SIZE,  // dwords of records following the header
  TH_PAD,
  TH_DWORDS
};
constexpr uint32_t TF_RESIDUAL = 1u;  // multiply by the residual (R0..R3)
constexpr uint32_t TF_PI = 2u;        // A or B may be non-zero

// A chunk is [n_lin, n_pair, base, n_pm] followed by its n_pm sign records,
// n_lin LIN records (mask, delta) and n_pair PAIR records; n_pm + n_lin +
// n_pair <= CHUNK_MAX.  Sign records are the LIN records whose packed delta
// is, mod 2^32, s 2^30 + t 2^9 with s, t in {-1, 0, 1} (S += 2s mod 8,
// N += t: every node-phase factor 1 + w^{4p + phi} with phi odd or 2 mod 4).
// They carry no delta: they come in groups of at most PM_GROUP, each a
// PM_HDR-dword header [S+, S-, N+, N-] and its records' masks.  The kernel
// shifts the group's parities into one word (record j of n at bit 32 - n + j)
// and adds 2^30 popcount((p & S+) | (~p & S-)) + 2^9 popcount((p & N+) |
// (~p & N-)); the base holds -2^30 |S-| - 2^9 |N-|.  Blobs of version 3 and 4
// have n_pm = 0 in every chunk.
constexpr int CHUNK_HDR = 4;  // n_lin, n_pair, base, n_pm | pad << PM_PAD_SHIFT
constexpr int PM_GROUP = 32;
constexpr int PM_HDR = 4;     // S+, S-, N+, N- masks over the group's records
inline uint64_t pm_dwords(uint64_t n_pm, uint64_t W32) {
  return (n_pm + PM_GROUP - 1) / PM_GROUP * PM_HDR + n_pm * W32;
}
// Parity groups (graphs flagged GF_TABLE, at most PM_TAB_W32_MAX mask words):
// a chunk's count word holds its number of groups (and, at bit
// PM_PAD_SHIFT, the count (< 4) of zero dwords between the chunk header and
// its first group, so that every group's tables start on a 16-byte boundary
// of the program words).  A group is the PM_TAB_HDR-dword header [S+, S-,
// N+, N-, X, Xc, J, Jc] and, for each of the 8 W32 nibbles j of the
// parameter words (bits 4j..4j+3), 16 dwords T_j[v], v < 16: bit b of
// T_j[v] is the parity of v & (nibble j of the mask of the form at bit b),
// so the XOR over j of T_j[nibble j of p] is the parity word y of the
// group's (at most 32) forms.  Forms are sign records (marked in S+/S- and
// N+/N- as above, counted in the chunk base the same way), PPX pairs (psi
// at bit 2m, phi at bit 2m + 1, X holding bit 2m, Xc their constant
// parities) and PROJ projectors (J holding the bit, Jc the parity the term
// needs).  A term of such a graph keeps no PROJ or PPX records of its own:
// they sit in the groups of its first chunk.
constexpr int PM_TAB_W32_MAX = 4;
// the kernels without the 128-bit path (8 waves per SIMD) hold a 1 KiB
// buffer per wave: tables of graphs of at most 2 mask words, in programs of
// at most PM_TAB_NARROW_WL lane words
constexpr int PM_TAB_W32_MAX_NARROW = 2;
constexpr int PM_TAB_NARROW_WL = 4;
constexpr int PM_TAB_HDR = 8;
constexpr uint32_t PM_COUNT_MASK = 0xffu, PM_PAD_SHIFT = 8;
constexpr uint32_t pm_tab_dwords(uint32_t W32) { return 128 * W32; }
constexpr uint32_t pm_tab_group_dwords(uint32_t W32) { return PM_TAB_HDR + pm_tab_dwords(W32); }

// graph header (dwords).  The terms of a graph are stored as GH_PARTS parts
// of about equal record work (the packer deals them out heaviest first), so
// that 1, 2 or 4 waves can split a graph's term sum: with L waves, wave j
// takes parts [j*GH_PARTS/L, (j+1)*GH_PARTS/L).  GH_PART_T[k] is the first
// term of part k (k = 1..3; part 0 starts at term 0) and GH_PART_O[k] its
// dword offset in the program.
constexpr int GH_PARTS = 4;
enum GraphHdr : int {
  GH_NTERMS = 0, GH_W32, GH_E0, GH_TERM_OFF, GH_NREC, GH_FLAGS,
  GH_PART_T1, GH_PART_T2, GH_PART_T3,
  GH_PART_O1, GH_PART_O2, GH_PART_O3,
  GH_APX_OFF,  // GF_APPROX: dword offset of the terms' complex64 factors (2 per term)
  GH_PAD1, GH_PAD2, GH_PAD3,
  GH_DWORDS
};
// graph flags
constexpr uint32_t GF_MONO = 1u;    // one term without pi powers or residual: its
                                    // float64 magnitude comes from a table
constexpr uint32_t GF_NARROW = 2u;  // every partial sum fits a signed 64-bit
                                    // integer (and no term uses several digits)
constexpr uint32_t GF_APPROX = 4u;  // approximate floatfactors: each term's exact
                                    // value is cast to float64, multiplied by its
                                    // complex64 factor and summed in float64, in
                                    // term order (no term split shares it)
constexpr uint32_t GF_STREAM = 8u;  // the 128-bit kernels read the graph's records
                                    // ahead into L2; the program then ends in
                                    // STREAM_PAD spare dwords
constexpr uint32_t GF_TABLE = 16u;  // sign groups carry parity tables instead of
                                    // masks (only in programs that run the
                                    // register-word kernels; layout.hpp's
                                    // PM_TAB_* limits)

// Record read-ahead of GF_STREAM graphs (the packer sets the flag on every
// graph of several terms): each touch reaches STREAM_AHEAD dwords past the
// exponent chunk about to be walked, so such a program ends in
// STREAM_PAD >= STREAM_AHEAD dwords past its last record.
constexpr int STREAM_AHEAD = 256;
constexpr int STREAM_PAD = 256;

// component header (dwords)
enum CompHdr : int {
  CH_NGRAPHS = 0, CH_F, CH_NOUT, CH_WMAX, CH_COLBASE, CH_FSEL_OFF,
  CH_DEST_OFF, CH_GRAPH_OFF,
  CH_FSEL0,  // first f column when f_selection is contiguous, else 0xffffffff
  CH_DEST0,  // first output column when dest is contiguous, else 0xffffffff
  CH_PAD1, CH_PAD2, CH_PAD3, CH_PAD4, CH_PAD5, CH_PAD6,
  CH_DWORDS
};

// program header (dwords); followed by n_comp component-header offsets
enum ProgHdr : int {
  PH_NCOMP = 0, PH_NUMF, PH_NUMOUT, PH_TOTOUT, PH_WMAX, PH_PAD0, PH_PAD1, PH_PAD2,
  PH_DWORDS
};

// device error bits
constexpr uint32_t ERR_TAB_RANGE = 1u;   // |n'| beyond the table
constexpr uint32_t ERR_PI_RANGE = 2u;    // pi exponents beyond PI_MAX
constexpr uint32_t ERR_SHIFT = 4u;       // alignment shift outside [0, max]
constexpr uint32_t ERR_WIDE = 8u;        // a 128-bit graph in a kernel built without that path
constexpr uint32_t ERR_FORMAT = 16u;     // parity tables in a kernel that reads only masks

// Largest alignment shift the 128-bit sum accepts; the packer proves
// bits(term) + shift + log2(G) <= 126 before upload.
constexpr int SHIFT_MAX = 126;

}  // namespace ksim
