// pack.hpp — host packer interface (FlatProgram graphs -> device words).
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/ksim_hip.h"
#include "layout.hpp"
#include "zomega.hpp"

namespace ksim {

struct PackStats {
  size_t graphs = 0, terms = 0, records = 0;
  int wmax = 1;         // 32-bit parameter words a lane holds (max over components)
  int n2_bound = 0;     // largest |n'| a term can reach
  size_t wide_terms = 0;  // terms that can reach beyond the single-digit rows
  int sum_bits = 0;     // bits of the 128-bit sum the worst graph may use
  double ops = 0;       // algorithmic 64-bit integer operations per shot
  size_t approx_graphs = 0;  // graphs with approximate floatfactors (GF_APPROX)
};

// Pack one graph: header into ghdr[GH_DWORDS], terms appended to out.
void pack_graph(const ksim_graph_view &g, std::vector<uint32_t> &out, uint32_t *ghdr,
                PackStats &st);

// Whole program; see ksim_program_create for the argument meaning.
std::vector<uint32_t> pack_program(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                                   const int32_t *f_selection, const int32_t *F_per_comp,
                                   int32_t n_comp, int32_t num_f, const int32_t *dest_col,
                                   int32_t num_outputs, PackStats &st);

const int32_t *materialise_table();  // TAB_DIGITS x TAB_N x 4 int32
const int32_t *pi_table();           // PI_MAX x PI_MAX x 4 int32

}  // namespace ksim
