// wire.hpp — the on-disk / on-wire form of a packed program.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "pack.hpp"

namespace ksim {

// Blob layout (little-endian), version 6 (16-dword graph headers with
// flags, term parts and the offset of a graph's approximate factors, graph
// i of a component ceil((F + i) / 32) mask words wide, up to LANE_WORDS_MAX
// words, record chunks that open with sign-record groups, held as parity
// tables in GF_TABLE graphs; version 5 had no parity tables, version 4 no
// sign records and version 3 neither approximate factors nor components
// wider than 8 words, and all three are read as well, every chunk of
// versions 3 and 4 holding zero sign records; version 2 let a graph without
// terms keep the width of its own parameter count, version 1 had 8-dword
// graph headers):
//   [0, 8)     magic "KSIMPRG\0"
//   [8, 12)    u32 version
//   [12, 16)   u32 header bytes (128)
//   [16, 24)   u64 n_words
//   [24, 32)   u64 FNV-1a 64 of the program words (as bytes)
//   [32, 96)   f64 x 8: the packer's stats (ksim_program_stats order)
//   [96, 116)  u32 x 5: n_comp, num_f, num_outputs, total outputs, lane words
//   [116, 128) zero
//   [128, ...) u32 program words
constexpr uint32_t WIRE_VERSION = 6;
constexpr size_t WIRE_HEADER = 128;

uint64_t fnv1a64(const void *data, size_t n);

std::vector<uint8_t> blob_from_words(const std::vector<uint32_t> &words, const double *stats8);

// Parse and fully validate a blob: every offset, count and size a kernel
// follows stays inside the words, so a corrupt or hostile blob is rejected
// here (std::invalid_argument) instead of reaching the device.
std::vector<uint32_t> words_from_blob(const uint8_t *blob, size_t n, double *stats8);

// The structural check alone (also run on freshly packed words in debug use).
void validate_words(const uint32_t *w, size_t n);

}  // namespace ksim
