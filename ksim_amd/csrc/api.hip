// api.hip — kernels, launchers and the C ABI of include/ksim_hip.h.
//
// Runtime layer of the sampler: owns the device copy of each packed program
// (the reference keeps four Kokkos Views per component,
// src/cpp/ksim/kokkos_sim.cpp:49-57,132-148), stages per-call inputs, launches
// the gfx950 kernels and reports device-side range errors.
#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ksim_hip.h"
#include "device.hpp"
#include "layout.hpp"
#include "pack.hpp"
#include "wire.hpp"

using namespace ksim;

namespace ksim {
__global__ void k_init_mag(const int4 *tab, double *mag);
}

namespace {

thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct RangeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F &&f) {
  try {
    f();
    return KSIM_OK;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return KSIM_E_INVALID;
  } catch (const RangeError &e) {
    g_last_error = e.what();
    return KSIM_E_RANGE;
  } catch (const DeviceError &e) {
    g_last_error = e.what();
    return KSIM_E_DEVICE;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return KSIM_E_UNSUPPORTED;
  }
}

// Per-device constant tables (materialisation and pi powers).
struct DevTables {
  int4 *tab = nullptr;
  int4 *pitab = nullptr;
  double *mag = nullptr;
  ksim::Tables view() const { return ksim::Tables{tab, pitab, mag}; }
};
std::mutex g_tab_mu;
std::map<int, DevTables> g_tabs;

DevTables tables_for_current_device() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_tab_mu);
  auto it = g_tabs.find(dev);
  if (it != g_tabs.end()) return it->second;
  DevTables t;
  hip_check(hipMalloc(&t.tab, sizeof(int4) * TAB_DIGITS * TAB_N), "hipMalloc(tab)");
  hip_check(hipMemcpy(t.tab, materialise_table(), sizeof(int4) * TAB_DIGITS * TAB_N,
                      hipMemcpyHostToDevice),
            "upload tab");
  hip_check(hipMalloc(&t.pitab, sizeof(int4) * PI_MAX * PI_MAX), "hipMalloc(pitab)");
  hip_check(hipMemcpy(t.pitab, pi_table(), sizeof(int4) * PI_MAX * PI_MAX,
                      hipMemcpyHostToDevice),
            "upload pitab");
  hip_check(hipMalloc(&t.mag, sizeof(double) * TAB_N), "hipMalloc(mag)");
  hipLaunchKernelGGL(ksim::k_init_mag, dim3((TAB_N + 255) / 256), dim3(256), 0, nullptr,
                     (const int4 *)t.tab, t.mag);
  hip_check(hipGetLastError(), "k_init_mag");
  hip_check(hipDeviceSynchronize(), "build magnitude table");
  g_tabs[dev] = t;
  return t;
}

std::string err_text(uint32_t e) {
  std::string s = "device exact-range check failed:";
  if (e & ERR_TAB_RANGE) s += " (1+sqrt2) exponent beyond the table";
  if (e & ERR_PI_RANGE) s += " pi exponent beyond the table";
  if (e & ERR_SHIFT) s += " alignment shift beyond the 128-bit sum";
  if (e & ERR_WIDE) s += " a graph needing 128-bit sums in a 64-bit kernel";
  if (e & ERR_FORMAT) s += " parity tables in a kernel that reads only masks";
  return s;
}

}  // namespace

// ------------------------------------------------------------------ kernels
namespace ksim {

constexpr int kBlock = 256;
// waves per SIMD the sampling kernels are compiled for: caps the VGPRs per
// lane (6 -> 80, 5 -> 96, 4 -> 128).  At 5 the hot loops hold no spills
// (at 6 the term loop spills to scratch), which measured fastest on the
// plain and the batched paths.
#ifndef KSIM_WAVES_PER_EU
#define KSIM_WAVES_PER_EU 5
#endif

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t *p, uint8_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// the magnitude table: magnitude() of every materialisation-table entry
__global__ void k_init_mag(const int4 *tab, double *mag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < TAB_N) mag[i] = table_magnitude(tab, i);
}

// Plain path: one lane per shot, grid (ceil(B/256), n_comp).  NOISE: the
// f bits are drawn in the prologue instead of read.  WIDE = false: a
// program whose graphs all have a single term or 64-bit sums (GF_MONO /
// GF_NARROW) runs without the 128-bit path, in fewer registers, so more
// waves per SIMD hide the scalar-load latency of its short graphs.
#ifndef KSIM_NARROW_WAVES_PER_EU
#define KSIM_NARROW_WAVES_PER_EU 8
#endif
template <int WL, bool NOISE = false, bool WIDE = true>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WIDE ? KSIM_WAVES_PER_EU : KSIM_NARROW_WAVES_PER_EU))) void k_sample(ShotArgs a, Tables tb) {
  __shared__ LdsTables lt;
  __shared__ uint32_t stage[4][NOISE ? 1 : stage_words(WL)];
  // the waves' parity-table buffers: 2 KiB in the 128-bit kernels, 1 KiB
  // in the others up to PM_TAB_NARROW_WL lane words
  constexpr int kTbl = WIDE ? 128 : (WL <= PM_TAB_NARROW_WL ? 64 : 0);
  __shared__ __attribute__((aligned(256))) uint4 tbl[kTbl ? 4 : 1][kTbl ? kTbl : 1];
  prefetch_program(a.prog, a.prefetch_dwords);
  load_tables(&lt, tb);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + 64 * wave;
  if (r0 >= a.B) return;  // the whole wave is past the last shot
  const int nrows = (int)min((int64_t)64, a.B - r0);
  const int64_t s = r0 + min(lane, nrows - 1);
  const FStage fs{NOISE ? nullptr : stage[wave], r0, nrows};
  const uint32_t tbuf = kTbl ? lds_addr(tbl[wave]) : NO_TBUF;
  run_shot<WL, NOISE, false, WIDE>(a, (int)blockIdx.y, s, &lt, tb,
                                   Split{1, 0, 0, nullptr, tbuf}, lane < nrows, nullptr, fs);
}

// Plain path with a term split: L waves (2 or 4) share each graph's terms
// on one 64-shot tile, for launches with too few shots to fill the GPU one
// lane per shot.  Block b takes the tile groups b, b + gridDim.x, ... (4/L
// tiles per group).
template <int WL, bool NOISE = false, bool WIDE = true>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WIDE ? KSIM_WAVES_PER_EU : KSIM_NARROW_WAVES_PER_EU))) void k_sample_split(
    ShotArgs a, Tables tb, int L) {
  __shared__ LdsTables lt;
  __shared__ __attribute__((aligned(256))) int4 xch[XCH_INT4];
  prefetch_program(a.prog, a.prefetch_dwords);
  load_tables(&lt, tb);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int G = 4 / L;
  const Split sp{L, wave % L, wave - wave % L, xch};
  const int64_t ntiles = (a.B + 63) >> 6;
  const int64_t ngroups = (ntiles + G - 1) / G;
  uint32_t xs = 0;  // exchange state, the same in every wave of the block
  for (int64_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
    // every wave of the block runs this iteration (the loop bound is the
    // block's), so the group's exchanges see all of its waves
    const int64_t tile = gi * G + wave / L;
    const int64_t off = tile * 64 + lane;
    const bool valid = off < a.B;
    const int64_t s = min(off, a.B - 1);
    run_shot<WL, NOISE, true, WIDE>(a, (int)blockIdx.y, s, &lt, tb, sp, valid, &xs);
  }
}

// Noise rows: lane per shot writes its f row (B, f_cols) u8, columns
// [0, num_f), 128 f bits per pass over the shot's noise stream.
__global__ __launch_bounds__(kBlock) void k_noise_rows(NoiseDev nz, int64_t B, int64_t shot_base,
                                                      int32_t num_f, int32_t f_cols,
                                                      uint8_t *__restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  uint8_t *row = out + s * (int64_t)f_cols;
  const bool dw = ((f_cols | (int32_t)(uintptr_t)out) & 3) == 0;
  for (int q = 0; 128 * q < num_f; ++q) {
    uint32_t buf[4] = {0u, 0u, 0u, 0u};
    if (nz.C > 0)
      noise_walk(nz, noise_stream(nz.seed, (uint64_t)(shot_base + s)), [&](int, uint32_t r) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (4 * q + k < nz.wf) buf[k] ^= nz.pat[r + 4 * q + k];
      });
    const int n = min(128, num_f - 128 * q);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      for (int j = 0; j < 32; j += 4) {
        const int col = 128 * q + 32 * k + j;
        if (32 * k + j >= n) break;
        const uint32_t b = (buf[k] >> j) & 0xfu;
        if (dw && 32 * k + j + 4 <= n) {
          *(uint32_t *)(row + col) = (b * 0x00204081u) & 0x01010101u;  // 4 bits -> 4 bytes
        } else {
          for (int t = 0; t < 4 && 32 * k + j + t < n; ++t) row[col + t] = (uint8_t)((b >> t) & 1u);
        }
      }
    }
  }
}

// Explicit parameter rows: exact value (mode 0) or |value| (mode 1) of one graph.
struct EvalArgs {
  const uint32_t *prog;
  const uint32_t *gh;
  const uint8_t *params;  // (B, P)
  int64_t B;
  int32_t P, mode;
  int64_t *coeffs;  // (B, 4, 2)
  int32_t *exp2;    // (B)
  double *absout;   // (B)
  uint32_t *err;
};

template <int WL>
__global__ __launch_bounds__(kBlock) void k_eval(EvalArgs a, Tables tb) {
  __shared__ LdsTables lt;
  __shared__ __attribute__((aligned(256))) uint4 tbl[4][128];
  load_tables(&lt, tb);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + 64 * wave;
  if (r0 >= a.B) return;  // the whole wave is past the last row
  // lanes past the last row evaluate a copy of it (the parity tables are
  // copied into LDS by all 64 lanes of a wave) and write nothing
  const int64_t s0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = s0 < a.B;
  const int64_t s = valid ? s0 : a.B - 1;
  const uint32_t tbuf = lds_addr(tbl[wave]);
  const uint8_t *row = a.params + s * (int64_t)a.P;
  uint32_t pv[WL];
#pragma unroll
  for (int w = 0; w < WL; ++w) {
    uint32_t word = 0;
    const int j0 = 32 * w, jn = min(32, a.P - j0);
    for (int jj = 0; jj < jn; ++jj) word |= (uint32_t)(row[j0 + jj] & 1u) << jj;
    pv[w] = word;
  }
  uint32_t err = 0;
  const int e0 = (int)a.gh[GH_E0];
  if (a.mode == 0) {
    Acc4 acc;
    eval_graph_any<WL>(as_const(a.prog), as_const(a.gh), pv, &lt, tb, acc, err, tbuf);
    if (valid) {
      int64_t *c = a.coeffs + s * 8;
      const i128_t v[4] = {acc.a, acc.b, acc.c, acc.d};
      for (int k = 0; k < 4; ++k) {
        c[2 * k] = (int64_t)(uint64_t)(u128_t)v[k];
        c[2 * k + 1] = (int64_t)(v[k] >> 64);
      }
      a.exp2[s] = e0;
    }
  } else {
    // the sampler's own probability path (table magnitudes, 64-bit sums)
    uint32_t xs = 0;
    const double m = graph_prob_any<WL>(as_const(a.prog), as_const(a.gh), pv, &lt, tb, err,
                                        Split{1, 0, 0, nullptr, tbuf}, xs);
    if (valid) a.absout[s] = m;
  }
  if (err) atomicOr(a.err, err);
}

// Batched persistent path: each block pulls work items from a device-wide
// cursor until the list is drained.  An item is shots [shot0, shot0 + count)
// of one component of one program with a term split L: the block's four
// waves form 4/L groups of L waves; group g takes the 64-shot tiles
// g, g + 4/L, ... of the item and its L waves share each graph's terms.
struct BatchProg {
  const uint32_t *prog;
  const uint8_t *f;
  uint8_t *out;
  uint64_t seed;
  int64_t B;
  int32_t num_f, num_outputs, tot_out, wmax;
};
struct BatchItem {
  int32_t prog, comp;
  int64_t shot0;
  int32_t count;  // shots [shot0, shot0 + count) of program `prog`
  int32_t split;  // waves per group: 1, 2 or 4
};

// reset the work cursor and the word `clear` (the batch clears its error
// word per launch; the scheduler passes a scratch word and keeps its error
// word sticky until checked)
__global__ void k_reset_cursor(uint32_t *cursor, uint32_t *clear) {
  if (threadIdx.x == 0) {
    cursor[0] = 0;
    clear[0] = 0;
  }
}

// The batch kernel stages f rows only where the extra LDS keeps the
// resident block count: the 128-bit kernels (5 blocks per CU) with at most
// 4 parameter words per lane.  The plain split kernel reads f lane by lane:
// its launches are too small for the f traffic to matter, and the staged
// read measured 3-21% slower there at L = 4 (more spilled registers).
template <int WL, bool WIDE>
constexpr bool kStageBatch = WIDE && WL <= 4;

template <int WL, bool WIDE = true>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WIDE ? KSIM_WAVES_PER_EU : KSIM_NARROW_WAVES_PER_EU))) void k_batch(
    const BatchProg *__restrict__ progs, const BatchItem *__restrict__ items, uint32_t n_items,
    uint32_t *cursor, uint32_t *err, Tables tb) {
  __shared__ LdsTables lt;
  __shared__ uint32_t s_item;
  __shared__ __attribute__((aligned(256))) int4 xch[XCH_INT4];
  __shared__ uint32_t stage[4][kStageBatch<WL, WIDE> ? stage_words(WL) : 1];
  load_tables(&lt, tb);
  uint32_t xs = 0;  // exchange state, the same in every wave of the block
  // the wave index is the same on every lane of a wave: readfirstlane lets
  // the compiler keep what derives from it (term ranges, program offsets)
  // on the scalar path
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(cursor, 1u);
    __syncthreads();
    const uint32_t idx = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();  // s_item is read before thread 0 takes the next item
    if (idx >= n_items) break;
    const BatchItem it = items[idx];
    const BatchProg bp = progs[it.prog];
    ShotArgs a;
    a.prog = bp.prog; a.f = bp.f; a.U = nullptr; a.out = bp.out; a.err = err;
    a.B = bp.B; a.shot_base = 0; a.seed = bp.seed;
    a.num_f = bp.num_f; a.num_outputs = bp.num_outputs; a.tot_out = bp.tot_out; a.prefetch_dwords = 0;
    a.out_words = 0; a.out_atomic = 0;
    a.noise = NoiseDev{};
    const int L = it.split, G = 4 / L;
    const Split sp{L, wave % L, wave - wave % L, xch};
    const int ntiles = (it.count + 63) >> 6;
    const int rounds = (ntiles + G - 1) / G;
    for (int r = 0; r < rounds; ++r) {
      const int tile = r * G + wave / L;
      // with L == 1 there are no barriers: a wave without a tile stops
      if (L == 1 && tile >= ntiles) break;
      const int off = tile * 64 + lane;
      const bool valid = off < it.count;
      const int64_t s = it.shot0 + min(off, it.count - 1);
      // a tile past the item's end (split groups) repeats its last shot
      const int t0 = min(tile * 64, it.count - 1);
      const FStage fs{kStageBatch<WL, WIDE> ? stage[wave] : nullptr, it.shot0 + t0,
                      min(64, it.count - t0)};
      run_shot<WL, false, true, WIDE>(a, it.comp, s, &lt, tb, sp, valid, &xs, fs);
    }
  }
}

// Components wider than the register words: each lane's parameter words in
// LDS (`dyn`, `words` dwords per lane, interleaved by lane within a wave),
// L waves (1, 2 or 4) per 64-shot tile; block b takes the tile groups
// b, b + gridDim.x, ... (4/L tiles per group).
template <bool NOISE, bool WIDE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WIDE ? KSIM_WAVES_PER_EU : KSIM_NARROW_WAVES_PER_EU))) void k_sample_lds(
    ShotArgs a, Tables tb, int L, int words) {
  __shared__ LdsTables lt;
  __shared__ int4 xch[XCH_INT4];
  extern __shared__ uint32_t dyn[];
  prefetch_program(a.prog, a.prefetch_dwords);
  load_tables(&lt, tb);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  uint32_t *pvl = dyn + (size_t)wave * words * 64 + lane;
  const int G = 4 / L;
  const Split sp{L, wave % L, wave - wave % L, xch};
  const int64_t ntiles = (a.B + 63) >> 6;
  const int64_t ngroups = (ntiles + G - 1) / G;
  uint32_t xs = 0;
  for (int64_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
    const int64_t tile = gi * G + wave / L;
    const int64_t off = tile * 64 + lane;
    const bool valid = off < a.B;
    const int64_t s = min(off, a.B - 1);
    run_shot_lds<NOISE, true, WIDE>(a, (int)blockIdx.y, s, &lt, tb, pvl, sp, valid, &xs);
  }
}

// k_eval for graphs wider than the register words (parameter words in LDS)
__global__ __launch_bounds__(kBlock) void k_eval_lds(EvalArgs a, Tables tb, int words) {
  __shared__ LdsTables lt;
  extern __shared__ uint32_t dyn[];
  load_tables(&lt, tb);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.B) return;
  uint32_t *pvl = dyn + (size_t)(threadIdx.x >> 6) * words * 64 + (threadIdx.x & 63);
  const uint8_t *row = a.params + s * (int64_t)a.P;
  for (int w = 0; w < words; ++w) {
    uint32_t word = 0;
    const int j0 = 32 * w, jn = min(32, a.P - j0);
    for (int jj = 0; jj < jn; ++jj) word |= (uint32_t)(row[j0 + jj] & 1u) << jj;
    pvl[64 * w] = word;
  }
  uint32_t err = 0;
  const cword *gh = as_const(a.gh);
  const LdsWords pw{pvl, (int)gh[GH_W32]};
  if (a.mode == 0) {
    Acc4 acc;
    eval_graph(as_const(a.prog), gh, pw, &lt, tb, acc, err, NO_TBUF);
    int64_t *c = a.coeffs + s * 8;
    const i128_t v[4] = {acc.a, acc.b, acc.c, acc.d};
    for (int k = 0; k < 4; ++k) {
      c[2 * k] = (int64_t)(uint64_t)(u128_t)v[k];
      c[2 * k + 1] = (int64_t)(v[k] >> 64);
    }
    a.exp2[s] = (int)gh[GH_E0];
  } else {
    uint32_t xs = 0;
    a.absout[s] = graph_prob<LdsWords, false, true>(as_const(a.prog), gh, pw, &lt, tb, err,
                                                    Split{1, 0, 0, nullptr}, xs);
  }
  if (err) atomicOr(a.err, err);
}

// k_batch for batches holding a component wider than the register words:
// every program's parameter words in LDS (`words` dwords per lane, the
// widest component's).
template <bool WIDE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WIDE ? KSIM_WAVES_PER_EU : KSIM_NARROW_WAVES_PER_EU))) void k_batch_lds(
    const BatchProg *__restrict__ progs, const BatchItem *__restrict__ items, uint32_t n_items,
    uint32_t *cursor, uint32_t *err, Tables tb, int words) {
  __shared__ LdsTables lt;
  __shared__ uint32_t s_item;
  __shared__ int4 xch[XCH_INT4];
  extern __shared__ uint32_t dyn[];
  load_tables(&lt, tb);
  uint32_t xs = 0;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  uint32_t *pvl = dyn + (size_t)wave * words * 64 + lane;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(cursor, 1u);
    __syncthreads();
    const uint32_t idx = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (idx >= n_items) break;
    const BatchItem it = items[idx];
    const BatchProg bp = progs[it.prog];
    ShotArgs a;
    a.prog = bp.prog; a.f = bp.f; a.U = nullptr; a.out = bp.out; a.err = err;
    a.B = bp.B; a.shot_base = 0; a.seed = bp.seed;
    a.num_f = bp.num_f; a.num_outputs = bp.num_outputs; a.tot_out = bp.tot_out;
    a.prefetch_dwords = 0; a.out_words = 0; a.out_atomic = 0;
    a.noise = NoiseDev{};
    const int L = it.split, G = 4 / L;
    const Split sp{L, wave % L, wave - wave % L, xch};
    const int ntiles = (it.count + 63) >> 6;
    const int rounds = (ntiles + G - 1) / G;
    for (int r = 0; r < rounds; ++r) {
      const int tile = r * G + wave / L;
      if (L == 1 && tile >= ntiles) break;
      const int off = tile * 64 + lane;
      const bool valid = off < it.count;
      const int64_t s = it.shot0 + min(off, it.count - 1);
      run_shot_lds<false, true, WIDE>(a, it.comp, s, &lt, tb, pvl, sp, valid, &xs);
    }
  }
}

}  // namespace ksim

// ------------------------------------------------------------------ objects
namespace ksim {
struct HostPipe;
void destroy_pipe(HostPipe *h);
}  // namespace ksim

struct ksim_program {
  int device = 0;
  uint32_t *d_prog = nullptr;
  size_t n_words = 0;
  uint32_t *d_err = nullptr;
  int n_comp = 0, num_f = 0, num_outputs = 0, tot_out = 0, wmax = 1, wl = 1;
  std::vector<int> n_out, F, n_graphs;
  std::vector<std::vector<uint32_t>> graph_hdr;  // host copy of graph headers per comp
  PackStats st;
  int block_threads = kBlock;
  int out_words = 0;              // packed output row words: ceil(num_outputs / 32)
  bool packed_full_rows = false;  // one component writes whole packed rows in order
  int term_split = 0;  // waves per 64-shot tile in plain launches; 0: chosen per launch
  bool all_narrow = false;  // every graph GF_MONO or GF_NARROW: kernels without 128-bit sums
  bool approx = false;      // some graph has approximate floatfactors (GF_APPROX)
  // staging for the host-pointer API
  uint8_t *d_f = nullptr;
  size_t f_cap = 0;
  uint8_t *d_out = nullptr;
  size_t out_cap = 0;
  double *d_U = nullptr;
  size_t U_cap = 0;
  uint8_t *d_params = nullptr;
  size_t params_cap = 0;
  void *d_res = nullptr;
  size_t res_cap = 0;
  HostPipe *pipe = nullptr;  // pinned chunk pipeline of the host-pointer API
};

namespace {

// Per-wave work of one shot of component c when L waves (1, 2, 4) share each
// graph's terms, in record units (a term costs about 40 records of
// materialisation, a graph about 60 of decision work; single-term graphs
// are evaluated by every wave of the group, the others add an exchange).
void comp_work(const ksim_program *p, int c, double w[3]) {
  w[0] = w[1] = w[2] = 0.0;
  const std::vector<uint32_t> &gh = p->graph_hdr[c];
  for (int i = 0; i < p->n_graphs[c]; ++i) {
    const uint32_t *h = &gh[(size_t)i * GH_DWORDS];
    const double body = (double)h[GH_NREC] + 40.0 * h[GH_NTERMS];
    w[0] += body + 60.0;
    // single-term and approximate graphs are evaluated whole by every wave
    const bool mono = h[GH_FLAGS] & (GF_MONO | GF_APPROX);
    w[1] += mono ? body + 60.0 : body / 2 + 80.0;
    w[2] += mono ? body + 60.0 : body / 4 + 80.0;
  }
}

// algorithmic ops per shot of component c (the program's ops shared out by work)
double comp_ops(const ksim_program *p, int c) {
  double tot = 0.0, mine = 0.0, w[3];
  for (int k = 0; k < p->n_comp; ++k) {
    comp_work(p, k, w);
    tot += w[0];
    if (k == c) mine = w[0];
  }
  return tot > 0 ? p->st.ops * mine / tot : 0.0;
}

// Term split for component c: the smallest L whose 64-shot tile costs each
// wave at most `tile_ops` algorithmic ops, among the splits that keep at
// least 60% of the waves' work useful; the cheapest such tile otherwise.
int split_for(const ksim_program *p, int c, double tile_ops) {
  double w[3];
  comp_work(p, c, w);
  if (w[0] <= 0) return 1;
  const double scale = comp_ops(p, c) / w[0];
  const int Ls[3] = {1, 2, 4};
  int best = 1;
  for (int k = 0; k < 3; ++k) {
    const double eff = w[0] / (Ls[k] * w[k]);
    if (k > 0 && eff < 0.6) break;
    best = Ls[k];
    if (64.0 * w[k] * scale <= tile_ops) break;
  }
  return best;
}

// per-wave algorithmic ops of a 64-shot tile of component c at split L
double tile_cost(const ksim_program *p, int c, int L) {
  double w[3];
  comp_work(p, c, w);
  if (w[0] <= 0) return 0.0;
  return 64.0 * w[L == 1 ? 0 : L == 2 ? 1 : 2] * comp_ops(p, c) / w[0];
}

}  // namespace

namespace {

// register words of the kernel that runs a program of wmax parameter words
// per lane: 1, 2, 4 or 8, or 0 for the kernels that keep them in LDS
int lane_words(int wmax) {
  return wmax <= 1 ? 1 : wmax <= 2 ? 2 : wmax <= 4 ? 4 : wmax <= REG_WORDS_MAX ? 8 : 0;
}

// dynamic LDS of the LDS-word kernels: four waves of `words` dwords per lane
size_t lds_words_bytes(int words) { return (size_t)4 * 64 * 4 * (size_t)words; }

// Make the LDS-word kernel `k` accept `bytes` of dynamic LDS, after checking
// that the device holds them beside the kernel's static LDS.
void allow_lds(const void *k, size_t bytes) {
  hipFuncAttributes fa;
  hip_check(hipFuncGetAttributes(&fa, k), "hipFuncGetAttributes");
  int dev = 0, max_block = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&max_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
            "LDS per block");
  if (fa.sharedSizeBytes + bytes > (size_t)max_block)
    throw std::invalid_argument("component too wide: its parameter words need " +
                                std::to_string(fa.sharedSizeBytes + bytes) +
                                " bytes of LDS per block, the device has " +
                                std::to_string(max_block));
  (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  (void)hipGetLastError();
}

template <class T>
void ensure(T *&p, size_t &cap, size_t bytes) {
  if (bytes <= cap && p) return;
  if (p) hip_check(hipFree(p), "hipFree");
  p = nullptr;
  cap = 0;
  hip_check(hipMalloc((void **)&p, std::max<size_t>(bytes, 64)), "hipMalloc(staging)");
  cap = std::max<size_t>(bytes, 64);
}

void fill_u8_async(uint8_t *p, uint8_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
  hip_check(hipGetLastError(), "k_fill_u8");
}

void reset_err_async(ksim_program *p, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, p->d_err, 0u, (int64_t)1);
  hip_check(hipGetLastError(), "k_fill_u32");
}

void check_err_sync(ksim_program *p, hipStream_t st) {
  uint32_t e = 0;
  hip_check(hipMemcpyAsync(&e, p->d_err, sizeof(e), hipMemcpyDeviceToHost, st), "read err");
  hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
  if (e) throw RangeError(err_text(e));
}

// program dwords a plain launch pulls into L2 before its waves start (a
// quarter of an XCD's 4 MiB L2)
constexpr size_t kPrefetchDwords = 256 * 1024;

// most blocks of a term-split plain launch (they loop over tile groups)
constexpr int kSplitBlocks = 2048;

// Resident 256-thread blocks per CU, i.e. waves per SIMD, of the plain
// kernel (split false) or the split kernel for lane words wl, 64-bit or
// 128-bit sums and with or without in-kernel noise: the runtime's occupancy
// figure for the compiled kernel, asked once per process.
int resident_waves(int wl, bool narrow, bool nz, bool split, int words = 0) {
  if (wl == 0) {
    // the LDS-word kernel, for this program's words (asked per launch size)
    const void *k = nz ? (narrow ? reinterpret_cast<const void *>(&k_sample_lds<true, false>)
                                 : reinterpret_cast<const void *>(&k_sample_lds<true, true>))
                       : (narrow ? reinterpret_cast<const void *>(&k_sample_lds<false, false>)
                                 : reinterpret_cast<const void *>(&k_sample_lds<false, true>));
    int nb = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kBlock, lds_words_bytes(words)),
              "occupancy");
    return std::max(nb, 1);
  }
  static const std::array<int, 32> occ = [] {
    std::array<int, 32> o{};
    auto q = [](const void *k) {
      int nb = 0;
      hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kBlock, 0), "occupancy");
      return std::max(nb, 1);
    };
#define KSIM_OCC(I, W, N, WD)                                                              \
  o[(I * 2 + !(WD)) * 4 + (N) * 2 + 0] = q(reinterpret_cast<const void *>(&k_sample<W, N, WD>)); \
  o[(I * 2 + !(WD)) * 4 + (N) * 2 + 1] =                                                  \
      q(reinterpret_cast<const void *>(&k_sample_split<W, N, WD>));
#define KSIM_OCC_W(I, W) \
  KSIM_OCC(I, W, false, false) KSIM_OCC(I, W, true, false) KSIM_OCC(I, W, false, true) KSIM_OCC(I, W, true, true)
    KSIM_OCC_W(0, 1) KSIM_OCC_W(1, 2) KSIM_OCC_W(2, 4) KSIM_OCC_W(3, 8)
#undef KSIM_OCC_W
#undef KSIM_OCC
    return o;
  }();
  const int i = wl <= 1 ? 0 : wl <= 2 ? 1 : wl <= 4 ? 2 : 3;
  return occ[(i * 2 + (narrow ? 1 : 0)) * 4 + (nz ? 2 : 0) + (split ? 1 : 0)];
}

// A SIMD whose last round holds fewer waves than this is latency bound: on
// 30-qubit programs (t = 24, 32) a round of 1-3 resident waves took about
// as long as 3.5 waves' worth of issue at full occupancy.
constexpr double kLatencyWaves = 3.5;

// Waves per 64-shot tile for a plain launch of B shots: p->term_split when
// set.  Otherwise the fill rule: 1 when the tiles of all components fill
// the resident wave slots, else the smallest L in {2, 4} that fills them among the splits
// whose estimated useful share of the waves' work is at least 3/4 (every
// component), or the largest such L (measured on 30-qubit programs at 100k
// shots: t = 8 is fastest at L = 2, t = 12..32 at L = 4).  The fill rule
// ignores how the waves fall into rounds of resident slots: L = 2 on
// 2600 tiles is two full rounds where L = 4 is three rounds of half the
// work each.  So every allowed L above the fill rule's is also priced as
// (waves per SIMD in whole rounds + the last round, at least
// kLatencyWaves) x group iterations per block x per-wave work, on the grid
// the launch would have, and one that is at least 20 % cheaper is
// taken instead (measured 15-27 % faster launches between 2600 and 6000
// tiles).  Only larger L: the work estimate overprices the exchanges of
// graphs with few terms, and letting it pick a smaller L made t = 12 and
// 16 at 100k shots 7-17 % slower.
int plain_split(const ksim_program *p, int64_t B, bool nz) {
  if (p->term_split > 0) return p->term_split;
  static const int simds = [] {
    int dev = 0, cus = 256;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    return cus * 4;
  }();
  const int slots = simds * KSIM_WAVES_PER_EU;
  const int64_t tiles = (B + 63) / 64;
  const int Ls[3] = {1, 2, 4};
  double w[3] = {0.0, 0.0, 0.0}, eff[3] = {1.0, 1.0, 1.0};
  for (int c = 0; c < p->n_comp; ++c) {
    double wc[3];
    comp_work(p, c, wc);
    for (int k = 0; k < 3; ++k) w[k] += wc[k];
    if (wc[0] > 0)
      for (int k = 1; k < 3; ++k) eff[k] = std::min(eff[k], wc[0] / (Ls[k] * wc[k]));
  }
  // the grid holds tiles x n_comp wave tiles (one grid row per component)
  const int64_t wave_tiles = tiles * p->n_comp;
  int fill = 1;
  if (wave_tiles < slots) {
    for (int k = 1; k < 3; ++k) {
      if (eff[k] < 0.75) break;
      fill = Ls[k];
      if (Ls[k] * wave_tiles >= slots) break;
    }
  }
  if (w[0] <= 0) return fill;
  // Price the launched grid: L = 1 runs one block per 4 tiles, L > 1 one
  // block per group of 4/L tiles, at most kSplitBlocks blocks per component
  // that loop over the groups.  Each block iteration (a unit: one group of
  // tiles of one component) puts one wave on each SIMD of its CU for w[k]
  // of work, so the SIMDs each take m = units / CUs waves, in rounds of o
  // resident waves, however the units are dealt to looping blocks (a block
  // that finishes its units frees its slot for a waiting one).
  const int cus = simds / 4;
  auto cost = [&](int k) {
    const int o = resident_waves(p->wl, p->all_narrow, nz, Ls[k] > 1, p->wmax);
    const int64_t per = 4 / Ls[k];  // tiles per block iteration
    const int64_t units = (tiles + per - 1) / per * p->n_comp;
    const int64_t m = (units + cus - 1) / cus;  // waves per SIMD
    const int64_t full = m / o, rem = m % o;
    return ((double)(full * o) + (rem ? std::max((double)rem, kLatencyWaves) : 0.0)) * w[k];
  };
  const int kf = fill == 1 ? 0 : fill == 2 ? 1 : 2;
  const double cf = cost(kf);
  int best = fill;
  double cb = cf;
  for (int k = kf + 1; k < 3; ++k) {
    if (eff[k] < 0.75) break;
    const double ck = cost(k);
    if (ck < 0.8 * cf && ck < cb) {
      best = Ls[k];
      cb = ck;
    }
  }
  return best;
}

void launch_sample(ksim_program *p, const uint8_t *d_f, int64_t B, int32_t f_cols,
                   int64_t shot_base, uint64_t seed, const double *d_U, uint8_t *d_out,
                   hipStream_t st, int out_words = 0, const NoiseDev *noise = nullptr) {
  if (!noise && f_cols < p->num_f)
    throw std::invalid_argument("f_params has fewer columns than f_selection needs");
  if (B <= 0 || p->n_comp == 0) return;
  DevTables t = tables_for_current_device();
  ShotArgs a;
  a.prog = p->d_prog; a.f = d_f; a.U = d_U; a.out = d_out; a.err = p->d_err;
  a.B = B; a.shot_base = shot_base; a.seed = seed;
  a.num_f = f_cols; a.num_outputs = p->num_outputs; a.tot_out = p->tot_out;
  a.prefetch_dwords = (uint32_t)std::min<size_t>(p->n_words, kPrefetchDwords);
  a.out_words = out_words;
  a.out_atomic = p->n_comp > 1 ? 1 : 0;
  a.noise = noise ? *noise : NoiseDev{};
  if (out_words > 0 && !p->packed_full_rows) {
    // rows assembled from several components or scattered bits: OR-ed into
    // zeroed words (a fill kernel, so the launch stays graph-capturable)
    const int64_t n = B * (int64_t)out_words;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       (uint32_t *)d_out, 0u, n);
  }
  const int64_t nb = (B + kBlock - 1) / kBlock;
  if (nb > 0x7fffffff) throw std::invalid_argument("too many shots for one launch");
  const bool nz = a.noise.C > 0;
  const int L = plain_split(p, B, nz);
  if (p->wl == 0) {
    // parameter words in LDS: one kernel for every L, blocks looping over
    // the tile groups
    const int64_t groups = ((B + 63) / 64 + (4 / L) - 1) / (4 / L);
    const int blocks = (int)std::min<int64_t>(groups, kSplitBlocks);
    dim3 g2((unsigned)blocks, (unsigned)p->n_comp);
    const size_t dyn = lds_words_bytes(p->wmax);
#define KSIM_LDS_LAUNCH(N, WD)                                                             \
  do {                                                                                     \
    allow_lds(reinterpret_cast<const void *>(&k_sample_lds<N, WD>), dyn);                  \
    hipLaunchKernelGGL((k_sample_lds<N, WD>), g2, dim3(kBlock), dyn, st, a, t.view(), L,    \
                       p->wmax);                                                           \
  } while (0)
    if (p->all_narrow) {
      if (nz) KSIM_LDS_LAUNCH(true, false);
      else KSIM_LDS_LAUNCH(false, false);
    } else {
      if (nz) KSIM_LDS_LAUNCH(true, true);
      else KSIM_LDS_LAUNCH(false, true);
    }
#undef KSIM_LDS_LAUNCH
    hip_check(hipGetLastError(), "k_sample_lds launch");
    return;
  }
  if (L > 1) {
    // blocks: one per group of 4/L tiles, at most kSplitBlocks; blocks
    // loop over the groups
    const int64_t groups = ((B + 63) / 64 + (4 / L) - 1) / (4 / L);
    const int blocks = (int)std::min<int64_t>(groups, kSplitBlocks);
    dim3 g2((unsigned)blocks, (unsigned)p->n_comp);
#define KSIM_SPLIT_LAUNCH(W, N, WD) \
  hipLaunchKernelGGL((k_sample_split<W, N, WD>), g2, dim3(kBlock), 0, st, a, t.view(), L)
#define KSIM_SPLIT_CASES(W)                                   \
  if (p->all_narrow) {                                        \
    if (nz) KSIM_SPLIT_LAUNCH(W, true, false);                \
    else KSIM_SPLIT_LAUNCH(W, false, false);                  \
  } else {                                                    \
    if (nz) KSIM_SPLIT_LAUNCH(W, true, true);                 \
    else KSIM_SPLIT_LAUNCH(W, false, true);                   \
  }
    switch (p->wl) {
      case 1: KSIM_SPLIT_CASES(1); break;
      case 2: KSIM_SPLIT_CASES(2); break;
      case 4: KSIM_SPLIT_CASES(4); break;
      default: KSIM_SPLIT_CASES(8); break;
    }
#undef KSIM_SPLIT_CASES
#undef KSIM_SPLIT_LAUNCH
    hip_check(hipGetLastError(), "k_sample_split launch");
    return;
  }
  dim3 grid((unsigned)nb, (unsigned)p->n_comp);
#define KSIM_SAMPLE_LAUNCH(W, N, WD) \
  hipLaunchKernelGGL((k_sample<W, N, WD>), grid, dim3(kBlock), 0, st, a, t.view())
#define KSIM_SAMPLE_CASES(W)                                  \
  if (p->all_narrow) {                                        \
    if (nz) KSIM_SAMPLE_LAUNCH(W, true, false);               \
    else KSIM_SAMPLE_LAUNCH(W, false, false);                 \
  } else {                                                    \
    if (nz) KSIM_SAMPLE_LAUNCH(W, true, true);                \
    else KSIM_SAMPLE_LAUNCH(W, false, true);                  \
  }
  switch (p->wl) {
    case 1: KSIM_SAMPLE_CASES(1); break;
    case 2: KSIM_SAMPLE_CASES(2); break;
    case 4: KSIM_SAMPLE_CASES(4); break;
    default: KSIM_SAMPLE_CASES(8); break;
  }
#undef KSIM_SAMPLE_CASES
#undef KSIM_SAMPLE_LAUNCH
  hip_check(hipGetLastError(), "k_sample launch");
}

// upload packed (and validated) words as a program object
ksim_program *upload_program(const std::vector<uint32_t> &words, const PackStats &st) {
  const int n_comp = (int)words[PH_NCOMP];
  const int num_f = (int)words[PH_NUMF];
  const int num_outputs = (int)words[PH_NUMOUT];
  auto *p = new ksim_program();
  try {
    hip_check(hipGetDevice(&p->device), "hipGetDevice");
    p->n_words = words.size();
    hip_check(hipMalloc(&p->d_prog, words.size() * 4), "hipMalloc(program)");
    hip_check(hipMemcpy(p->d_prog, words.data(), words.size() * 4, hipMemcpyHostToDevice),
              "upload program");
    hip_check(hipMalloc(&p->d_err, 64), "hipMalloc(err)");
    hip_check(hipMemset(p->d_err, 0, 64), "hipMemset(err)");
    p->n_comp = n_comp;
    p->num_f = num_f;
    p->num_outputs = num_outputs;
    p->tot_out = (int)words[PH_TOTOUT];
    p->wmax = (int)words[PH_WMAX];
    p->wl = lane_words(p->wmax);
    p->out_words = (num_outputs + 31) / 32;
    if (n_comp == 1) {
      const uint32_t *c0 = &words[words[PH_DWORDS]];
      p->packed_full_rows = c0[CH_DEST0] == 0 && (int)c0[CH_NOUT] == num_outputs;
    }
    p->st = st;
    for (int c = 0; c < n_comp; ++c) {
      const uint32_t *ch = &words[words[PH_DWORDS + c]];
      p->n_graphs.push_back((int)ch[CH_NGRAPHS]);
      p->n_out.push_back((int)ch[CH_NOUT]);
      p->F.push_back((int)ch[CH_F]);
      std::vector<uint32_t> gh(&words[ch[CH_GRAPH_OFF]],
                               &words[ch[CH_GRAPH_OFF]] + (size_t)ch[CH_NGRAPHS] * GH_DWORDS);
      p->graph_hdr.push_back(std::move(gh));
    }
    p->all_narrow = true;
    for (const auto &gh : p->graph_hdr)
      for (size_t i = 0; i + GH_DWORDS <= gh.size(); i += GH_DWORDS) {
        if (!(gh[i + GH_FLAGS] & (GF_MONO | GF_NARROW))) p->all_narrow = false;
        if (gh[i + GH_FLAGS] & GF_APPROX) p->approx = true;
      }
    // approximate graphs have their float path only in the LDS-word kernels
    if (p->approx) p->wl = 0;
    tables_for_current_device();
  } catch (...) {
    ksim_program_destroy(p);
    throw;
  }
  return p;
}

ksim_program *create_program(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                             const int32_t *f_selection, const int32_t *F_per_comp,
                             int32_t n_comp, int32_t num_f, const int32_t *dest_col,
                             int32_t num_outputs) {
  if (!graphs && n_comp > 0) throw std::invalid_argument("graphs is NULL");
  PackStats st;
  std::vector<uint32_t> words = pack_program(graphs, graphs_per_comp, f_selection, F_per_comp,
                                             n_comp, num_f, dest_col, num_outputs, st);
  return upload_program(words, st);
}

void stats_to8(const PackStats &st, size_t n_words, double *s8) {
  const double v[8] = {(double)st.graphs, (double)st.terms, (double)st.records, (double)n_words,
                       st.ops, (double)st.n2_bound, (double)st.wide_terms, (double)st.sum_bits};
  std::memcpy(s8, v, sizeof(v));
}

PackStats stats_from8(const double *s8, int wmax) {
  PackStats st;
  st.graphs = (size_t)s8[0]; st.terms = (size_t)s8[1]; st.records = (size_t)s8[2];
  st.ops = s8[4]; st.n2_bound = (int)s8[5]; st.wide_terms = (size_t)s8[6];
  st.sum_bits = (int)s8[7]; st.wmax = wmax;
  return st;
}

void copy_out(const std::vector<uint8_t> &b, uint8_t *out, int64_t cap, int64_t *n_bytes) {
  if (!n_bytes) throw std::invalid_argument("n_bytes is NULL");
  *n_bytes = (int64_t)b.size();
  if (out && cap > 0) std::memcpy(out, b.data(), std::min<size_t>(b.size(), (size_t)cap));
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char *ksim_last_error(void) { return g_last_error.c_str(); }

int ksim_program_create(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                        const int32_t *f_selection, const int32_t *F_per_comp, int32_t n_comp,
                        int32_t num_f, const int32_t *dest_col, int32_t num_outputs,
                        ksim_program **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    *out = create_program(graphs, graphs_per_comp, f_selection, F_per_comp, n_comp, num_f,
                          dest_col, num_outputs);
  });
}

int ksim_component_create(const ksim_graph_view *graphs, int32_t n_graphs, int32_t num_f,
                          ksim_program **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (n_graphs < 1) throw std::invalid_argument("a component needs at least one graph");
    if (num_f < 0) throw std::invalid_argument("num_f < 0");
    std::vector<int32_t> fsel(num_f);
    for (int j = 0; j < num_f; ++j) fsel[j] = j;
    const int n_out = n_graphs - 1;
    std::vector<int32_t> dest(std::max(n_out, 0));
    for (int j = 0; j < n_out; ++j) dest[j] = j;
    const int32_t gpc = n_graphs, F = num_f;
    *out = create_program(graphs, &gpc, fsel.data(), &F, 1, num_f, dest.data(),
                          std::max(n_out, 1));
  });
}

int32_t ksim_component_n_out(const ksim_program *p) {
  return (p && p->n_comp > 0) ? p->n_out[0] : -1;
}

int ksim_program_sample_device(ksim_program *p, const uint8_t *d_f, int64_t B, int32_t f_cols,
                               int64_t shot_base, uint64_t seed, const double *d_U,
                               uint8_t *d_out, void *stream) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (B < 0) throw std::invalid_argument("B < 0");
    if (B > 0 && ((!d_f && f_cols > 0) || !d_out)) throw std::invalid_argument("NULL buffer");
    hipStream_t st = (hipStream_t)stream;
    reset_err_async(p, st);
    launch_sample(p, d_f, B, f_cols, shot_base, seed, d_U, d_out, st);
  });
}

int ksim_program_sample_device_packed(ksim_program *p, const uint8_t *d_f, int64_t B,
                                      int32_t f_cols, int64_t shot_base, uint64_t seed,
                                      const double *d_U, uint32_t *d_out, void *stream) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (B < 0) throw std::invalid_argument("B < 0");
    if (B > 0 && ((!d_f && f_cols > 0) || !d_out)) throw std::invalid_argument("NULL buffer");
    if (p->out_words == 0) return;
    hipStream_t st = (hipStream_t)stream;
    reset_err_async(p, st);
    launch_sample(p, d_f, B, f_cols, shot_base, seed, d_U, (uint8_t *)d_out, st, p->out_words);
  });
}

int ksim_program_check(ksim_program *p, void *stream) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    check_err_sync(p, (hipStream_t)stream);
  });
}

}  // extern "C"

namespace ksim {

// Copies between caller memory and pinned staging, split over a few
// persistent threads in pieces of at least 64 KiB: one thread's memcpy
// moves a fraction of what the host memory and the PCIe link can, and a
// fresh output array takes a page fault per 4 KiB on first touch, which
// threads take in parallel (a 29 MB copy into a new numpy array: about
// 8 ms on one thread, 2 ms on eight).
class CopyPool {
 public:
  static CopyPool &get() {
    static CopyPool pool;
    return pool;
  }
  void copy(void *dst, const void *src, size_t n) {
    const size_t parts = std::min<size_t>(workers_.size() + 1, std::max<size_t>(1, n >> 16));
    if (parts <= 1) {
      std::memcpy(dst, src, n);
      return;
    }
    const size_t piece = (n + parts - 1) / parts;
    int mine = 0;
    {
      std::lock_guard<std::mutex> lk(m_);
      for (size_t i = 1; i < parts && i * piece < n; ++i) {
        jobs_.push_back(Job{(char *)dst + i * piece, (const char *)src + i * piece,
                            std::min(piece, n - i * piece), &mine});
        ++mine;
      }
    }
    cv_.notify_all();
    std::memcpy(dst, src, std::min(piece, n));
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [&] { return mine == 0; });
  }

 private:
  struct Job {
    char *d;
    const char *s;
    size_t n;
    int *left;  // the caller's count of pieces not yet copied
  };
  CopyPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nt = std::min(7u, hw > 1 ? hw - 1 : 0u);
    for (unsigned i = 0; i < nt; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : workers_) t.join();
  }
  void loop() {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
        if (jobs_.empty()) return;
        j = jobs_.front();
        jobs_.pop_front();
      }
      std::memcpy(j.d, j.s, j.n);
      {
        std::lock_guard<std::mutex> lk(m_);
        --*j.left;
      }
      done_.notify_all();
    }
  }
  std::mutex m_;
  std::condition_variable cv_, done_;
  std::deque<Job> jobs_;
  bool stop_ = false;
  std::vector<std::thread> workers_;
};

// Host-pointer sampling of large inputs as a pipeline of chunks of whole
// 64-shot tiles: chunk k's f (and U) rows are copied into a pinned slot and
// sent on the upload stream while chunk k-1 runs on the compute stream and
// chunk k-2's rows come back on the download stream.  Two slots of each.
struct HostPipe {
  static constexpr int S = 2;
  hipStream_t up = nullptr, run = nullptr, down = nullptr;
  hipEvent_t sent[S] = {}, ran[S] = {}, back[S] = {};
  uint8_t *pin_f[S] = {}, *pin_o[S] = {};
  double *pin_u[S] = {};
  size_t cap_f = 0, cap_o = 0, cap_u = 0;
};

void destroy_pipe(HostPipe *h) {
  if (!h) return;
  for (int s = 0; s < HostPipe::S; ++s) {
    if (h->sent[s]) (void)hipEventDestroy(h->sent[s]);
    if (h->ran[s]) (void)hipEventDestroy(h->ran[s]);
    if (h->back[s]) (void)hipEventDestroy(h->back[s]);
    if (h->pin_f[s]) (void)hipHostFree(h->pin_f[s]);
    if (h->pin_o[s]) (void)hipHostFree(h->pin_o[s]);
    if (h->pin_u[s]) (void)hipHostFree(h->pin_u[s]);
  }
  if (h->up) (void)hipStreamDestroy(h->up);
  if (h->run) (void)hipStreamDestroy(h->run);
  if (h->down) (void)hipStreamDestroy(h->down);
  delete h;
}

}  // namespace ksim

namespace {

// bytes of f and U rows per pipeline chunk, and the input size from which
// the host path pipelines
constexpr size_t kChunkBytes = 8u << 20;
constexpr size_t kPipeMinBytes = 2 * kChunkBytes;

HostPipe *pipe_for(ksim_program *p, size_t fb, size_t ob, size_t ub) {
  if (!p->pipe) {
    std::unique_ptr<HostPipe> h(new HostPipe);
    for (hipStream_t *q : {&h->up, &h->run, &h->down})
      hip_check(hipStreamCreateWithFlags(q, hipStreamNonBlocking), "hipStreamCreate");
    for (int s = 0; s < HostPipe::S; ++s)
      for (hipEvent_t *e : {&h->sent[s], &h->ran[s], &h->back[s]})
        hip_check(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    p->pipe = h.release();
  }
  HostPipe *h = p->pipe;
  auto grow = [](auto **slot, size_t &cap, size_t need) {
    if (need <= cap) return;
    for (int s = 0; s < HostPipe::S; ++s) {
      if (slot[s]) hip_check(hipHostFree(slot[s]), "hipHostFree");
      slot[s] = nullptr;
    }
    cap = 0;
    for (int s = 0; s < HostPipe::S; ++s)
      hip_check(hipHostMalloc((void **)&slot[s], need, hipHostMallocDefault), "hipHostMalloc");
    cap = need;
  };
  grow(h->pin_f, h->cap_f, std::max<size_t>(fb, 64));
  grow(h->pin_o, h->cap_o, std::max<size_t>(ob, 64));
  if (ub) grow(h->pin_u, h->cap_u, ub);
  return h;
}

// Map the pages of a caller's output array writable ahead of the copies
// into it, on a thread of its own: a fresh array faults on every 4 KiB page
// at first touch, which otherwise stalls the copies on the pipeline's
// critical path (MADV_POPULATE_WRITE leaves the contents as they are, so
// it may run beside the copies; kernels without it skip this).
class Populate {
 public:
  Populate(void *p, size_t n) {
    const uintptr_t pg = (uintptr_t)sysconf(_SC_PAGESIZE);
    const uintptr_t a = ((uintptr_t)p + pg - 1) & ~(pg - 1);
    const uintptr_t b = ((uintptr_t)p + n) & ~(pg - 1);
    if (b > a) th_ = std::thread([a, b] {
        constexpr int kPopulateWrite = 23;  // MADV_POPULATE_WRITE (Linux 5.14)
        for (uintptr_t x = a; x < b; x += (4u << 20))
          if (madvise((void *)x, std::min<uintptr_t>(4u << 20, b - x), kPopulateWrite) != 0)
            return;
      });
  }
  ~Populate() {
    if (th_.joinable()) th_.join();
  }

 private:
  std::thread th_;
};

// The chunked host path (same results as one launch: every shot draws the
// stream keyed by its index in the whole call, `shot_base`).
void sample_host_pipelined(ksim_program *p, const uint8_t *f, int64_t B, int32_t f_cols,
                           uint64_t seed, const double *U, uint8_t *out, int out_cols,
                           bool packed) {
  const size_t frow = (size_t)f_cols, urow = U ? (size_t)p->tot_out * sizeof(double) : 0;
  const size_t orow = packed ? (size_t)p->out_words * 4 : (size_t)out_cols;
  const int64_t rows = std::max<int64_t>(64, (int64_t)(kChunkBytes / std::max<size_t>(frow + urow, 1)) & ~63);
  const int64_t nch = (B + rows - 1) / rows;
  ensure(p->d_f, p->f_cap, (size_t)B * frow);
  ensure(p->d_out, p->out_cap, (size_t)B * orow);
  if (U) ensure(p->d_U, p->U_cap, (size_t)B * urow);
  HostPipe *h = pipe_for(p, (size_t)rows * frow, (size_t)rows * orow, (size_t)rows * urow);
  CopyPool &cp = CopyPool::get();
  Populate pop(out, (size_t)B * orow);
  const bool fill = !packed && (out_cols != p->num_outputs || p->tot_out < p->num_outputs);
  auto take_back = [&](int64_t k) {  // chunk k's rows to the caller
    const int s = (int)(k % HostPipe::S);
    const int64_t r0 = k * rows, n = std::min(rows, B - r0);
    hip_check(hipEventSynchronize(h->back[s]), "hipEventSynchronize");
    cp.copy(out + (size_t)r0 * orow, h->pin_o[s], (size_t)n * orow);
  };
  reset_err_async(p, h->run);
  for (int64_t k = 0; k < nch; ++k) {
    const int s = (int)(k % HostPipe::S);
    const int64_t r0 = k * rows, n = std::min(rows, B - r0);
    // the slot's previous upload is over before its pinned rows are rewritten
    if (k >= HostPipe::S) hip_check(hipEventSynchronize(h->sent[s]), "hipEventSynchronize");
    if (frow) {
      cp.copy(h->pin_f[s], f + (size_t)r0 * frow, (size_t)n * frow);
      hip_check(hipMemcpyAsync(p->d_f + (size_t)r0 * frow, h->pin_f[s], (size_t)n * frow,
                               hipMemcpyHostToDevice, h->up), "H2D f");
    }
    if (urow) {
      cp.copy(h->pin_u[s], (const uint8_t *)U + (size_t)r0 * urow, (size_t)n * urow);
      hip_check(hipMemcpyAsync((uint8_t *)p->d_U + (size_t)r0 * urow, h->pin_u[s],
                               (size_t)n * urow, hipMemcpyHostToDevice, h->up), "H2D U");
    }
    hip_check(hipEventRecord(h->sent[s], h->up), "hipEventRecord");
    hip_check(hipStreamWaitEvent(h->run, h->sent[s], 0), "hipStreamWaitEvent");
    uint8_t *d_out = p->d_out + (size_t)r0 * orow;
    if (fill) fill_u8_async(d_out, 0, n * (int64_t)orow, h->run);
    launch_sample(p, p->d_f + (size_t)r0 * frow, n, f_cols, r0, seed,
                  U ? p->d_U + (size_t)r0 * p->tot_out : nullptr, d_out, h->run,
                  packed ? p->out_words : 0);
    hip_check(hipEventRecord(h->ran[s], h->run), "hipEventRecord");
    // the slot's previous rows reach the caller before it is reused
    if (k >= HostPipe::S) take_back(k - HostPipe::S);
    hip_check(hipStreamWaitEvent(h->down, h->ran[s], 0), "hipStreamWaitEvent");
    hip_check(hipMemcpyAsync(h->pin_o[s], d_out, (size_t)n * orow, hipMemcpyDeviceToHost,
                             h->down), "D2H out");
    hip_check(hipEventRecord(h->back[s], h->down), "hipEventRecord");
  }
  for (int64_t k = std::max<int64_t>(0, nch - HostPipe::S); k < nch; ++k) take_back(k);
  check_err_sync(p, h->run);
}

}  // namespace

extern "C" {

static void sample_host(ksim_program *p, const uint8_t *f, int64_t B, int32_t f_cols,
                        uint64_t seed, const double *U, uint8_t *out, int out_cols,
                        bool packed = false) {
  if (B < 0) throw std::invalid_argument("B < 0");
  if (f_cols < p->num_f) throw std::invalid_argument("f_params has fewer columns than f_selection needs");
  if (B == 0) return;
  if (!out || (!f && f_cols > 0)) throw std::invalid_argument("NULL buffer");
  hipStream_t st = nullptr;
  const size_t fb = (size_t)B * f_cols;
  const size_t ob = packed ? (size_t)B * p->out_words * 4 : (size_t)B * out_cols;
  const size_t ub = U ? (size_t)B * p->tot_out * sizeof(double) : 0;
  if (fb + ub >= kPipeMinBytes && B >= 256) {
    sample_host_pipelined(p, f, B, f_cols, seed, U, out, out_cols, packed);
    return;
  }
  ensure(p->d_f, p->f_cap, fb);
  ensure(p->d_out, p->out_cap, ob);
  if (fb) hip_check(hipMemcpyAsync(p->d_f, f, fb, hipMemcpyHostToDevice, st), "H2D f");
  const double *dU = nullptr;
  if (U) {
    const size_t ub = (size_t)B * p->tot_out * sizeof(double);
    ensure(p->d_U, p->U_cap, ub);
    if (ub) hip_check(hipMemcpyAsync(p->d_U, U, ub, hipMemcpyHostToDevice, st), "H2D U");
    dU = p->d_U;
  }
  if (!packed && (out_cols != p->num_outputs || p->tot_out < p->num_outputs))
    fill_u8_async(p->d_out, 0, (int64_t)ob, st);
  reset_err_async(p, st);
  launch_sample(p, p->d_f, B, f_cols, 0, seed, dU, p->d_out, st, packed ? p->out_words : 0);
  hip_check(hipMemcpyAsync(out, p->d_out, ob, hipMemcpyDeviceToHost, st), "D2H out");
  check_err_sync(p, st);
}

int ksim_program_sample_packed(ksim_program *p, const uint8_t *f_params, int64_t B,
                               int32_t f_cols, uint64_t seed, uint32_t *out) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (p->out_words == 0) return;
    sample_host(p, f_params, B, f_cols, seed, nullptr, (uint8_t *)out, 0, true);
  });
}

int ksim_program_sample(ksim_program *p, const uint8_t *f_params, int64_t B, int32_t f_cols,
                        uint64_t seed, uint8_t *out) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    sample_host(p, f_params, B, f_cols, seed, nullptr, out, p->num_outputs);
  });
}

int ksim_program_sample_uniforms(ksim_program *p, const uint8_t *f_params, int64_t B,
                                 int32_t f_cols, const double *U, uint8_t *out) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (!U && p->tot_out > 0 && B > 0) throw std::invalid_argument("U is NULL");
    sample_host(p, f_params, B, f_cols, 0, U, out, p->num_outputs);
  });
}

int ksim_component_sample(ksim_program *p, const uint8_t *f_sel, int64_t B, int32_t f_cols,
                          uint64_t seed, uint8_t *out) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (f_cols != p->num_f) throw std::invalid_argument("f_sel column count != F");
    sample_host(p, f_sel, B, f_cols, seed, nullptr, out, p->num_outputs);
  });
}

static void eval_host(ksim_program *p, int32_t comp, int32_t graph, const uint8_t *params,
                      int64_t B, int32_t P, int mode, int64_t *coeffs, int32_t *exp2,
                      double *absout) {
  if (comp < 0 || comp >= p->n_comp) throw std::invalid_argument("component index out of range");
  if (graph < 0 || graph >= p->n_graphs[comp]) throw std::invalid_argument("graph index out of range");
  const int expectP = p->F[comp] + graph;
  if (P != expectP) throw std::invalid_argument("params column count != P of the graph");
  if (B < 0) throw std::invalid_argument("B < 0");
  if (B == 0) return;
  const int W32 = std::max(1, (P + 31) / 32);
  const bool approx = p->graph_hdr[comp][(size_t)graph * GH_DWORDS + GH_FLAGS] & GF_APPROX;
  if (approx && mode == 0)
    throw std::runtime_error("the graph has approximate floatfactors: its value is not exact");
  const int wl = approx ? 0 : lane_words(W32);  // 0: parameter words in LDS
  hipStream_t st = nullptr;
  ensure(p->d_params, p->params_cap, (size_t)B * P);
  if ((size_t)B * P)
    hip_check(hipMemcpyAsync(p->d_params, params, (size_t)B * P, hipMemcpyHostToDevice, st),
              "H2D params");
  const size_t rb = mode == 0 ? (size_t)B * (8 * 8 + 4) : (size_t)B * 8;
  ensure(p->d_res, p->res_cap, rb);
  DevTables t = tables_for_current_device();
  // graph header lives in the device program: offset of component header + graph index
  uint32_t ch_off = 0;
  hip_check(hipMemcpy(&ch_off, p->d_prog + PH_DWORDS + comp, 4, hipMemcpyDeviceToHost),
            "read comp offset");
  uint32_t chd[CH_DWORDS];
  hip_check(hipMemcpy(chd, p->d_prog + ch_off, sizeof(chd), hipMemcpyDeviceToHost),
            "read comp header");
  EvalArgs a;
  a.prog = p->d_prog;
  a.gh = p->d_prog + chd[CH_GRAPH_OFF] + (size_t)graph * GH_DWORDS;
  a.params = p->d_params;
  a.B = B; a.P = P; a.mode = mode;
  a.coeffs = (int64_t *)p->d_res;
  a.exp2 = (int32_t *)((char *)p->d_res + (size_t)B * 64);
  a.absout = (double *)p->d_res;
  a.err = p->d_err;
  reset_err_async(p, st);
  dim3 grid((unsigned)((B + kBlock - 1) / kBlock));
  switch (wl) {
    case 1: hipLaunchKernelGGL(k_eval<1>, grid, dim3(kBlock), 0, st, a, t.view()); break;
    case 2: hipLaunchKernelGGL(k_eval<2>, grid, dim3(kBlock), 0, st, a, t.view()); break;
    case 4: hipLaunchKernelGGL(k_eval<4>, grid, dim3(kBlock), 0, st, a, t.view()); break;
    case 8: hipLaunchKernelGGL(k_eval<8>, grid, dim3(kBlock), 0, st, a, t.view()); break;
    default: {
      const size_t dyn = lds_words_bytes(W32);
      allow_lds(reinterpret_cast<const void *>(&k_eval_lds), dyn);
      hipLaunchKernelGGL(k_eval_lds, grid, dim3(kBlock), dyn, st, a, t.view(), W32);
      break;
    }
  }
  hip_check(hipGetLastError(), "k_eval launch");
  if (mode == 0) {
    hip_check(hipMemcpyAsync(coeffs, a.coeffs, (size_t)B * 64, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(exp2, a.exp2, (size_t)B * 4, hipMemcpyDeviceToHost, st), "D2H");
  } else {
    hip_check(hipMemcpyAsync(absout, a.absout, (size_t)B * 8, hipMemcpyDeviceToHost, st), "D2H");
  }
  check_err_sync(p, st);
}

int ksim_program_eval_exact(ksim_program *p, int32_t comp, int32_t graph,
                            const uint8_t *params, int64_t B, int32_t P, int64_t *coeffs,
                            int32_t *exp2) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (B > 0 && (!coeffs || !exp2 || (!params && P > 0))) throw std::invalid_argument("NULL buffer");
    eval_host(p, comp, graph, params, B, P, 0, coeffs, exp2, nullptr);
  });
}

int ksim_component_eval_abs(ksim_program *p, int32_t graph, const uint8_t *params, int64_t B,
                            int32_t P, double *out) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (B > 0 && (!out || (!params && P > 0))) throw std::invalid_argument("NULL buffer");
    eval_host(p, 0, graph, params, B, P, 1, nullptr, nullptr, out);
  });
}

int ksim_program_set_tiles(ksim_program *p, int32_t block_threads, int32_t term_split) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    if (block_threads != kBlock) throw std::invalid_argument("block_threads must be 256");
    if (term_split != 0 && term_split != 1 && term_split != 2 && term_split != 4)
      throw std::invalid_argument("term_split must be 0 (automatic), 1, 2 or 4");
    p->block_threads = block_threads;
    p->term_split = term_split;
  });
}

int32_t ksim_program_exact(const ksim_program *p) {
  return p ? (p->approx ? 0 : 1) : -1;
}

int64_t ksim_program_device_bytes(const ksim_program *p) {
  return p ? (int64_t)(p->n_words * 4) : -1;
}

int ksim_program_stats(const ksim_program *p, double *stats8) {
  return guarded([&] {
    if (!p || !stats8) throw std::invalid_argument("NULL argument");
    stats8[0] = (double)p->st.graphs;
    stats8[1] = (double)p->st.terms;
    stats8[2] = (double)p->st.records;
    stats8[3] = (double)p->n_words;
    stats8[4] = p->st.ops;
    stats8[5] = (double)p->st.n2_bound;
    stats8[6] = (double)p->st.wide_terms;
    stats8[7] = (double)p->st.sum_bits;
  });
}

void ksim_program_destroy(ksim_program *p) {
  if (!p) return;
  if (p->d_prog) (void)hipFree(p->d_prog);
  if (p->d_err) (void)hipFree(p->d_err);
  if (p->d_f) (void)hipFree(p->d_f);
  if (p->d_out) (void)hipFree(p->d_out);
  if (p->d_U) (void)hipFree(p->d_U);
  if (p->d_params) (void)hipFree(p->d_params);
  if (p->d_res) (void)hipFree(p->d_res);
  destroy_pipe(p->pipe);
  delete p;
}

int ksim_program_pack(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                      const int32_t *f_selection, const int32_t *F_per_comp, int32_t n_comp,
                      int32_t num_f, const int32_t *dest_col, int32_t num_outputs,
                      uint32_t *words, int64_t cap, int64_t *n_words, double *stats8) {
  return guarded([&] {
    if (!n_words) throw std::invalid_argument("n_words is NULL");
    PackStats st;
    std::vector<uint32_t> w = pack_program(graphs, graphs_per_comp, f_selection, F_per_comp,
                                           n_comp, num_f, dest_col, num_outputs, st);
    *n_words = (int64_t)w.size();
    if (stats8) {
      const double v[8] = {(double)st.graphs, (double)st.terms, (double)st.records,
                           (double)w.size(), st.ops, (double)st.n2_bound,
                           (double)st.wide_terms, (double)st.sum_bits};
      std::memcpy(stats8, v, sizeof(v));
    }
    if (words && cap > 0)
      std::memcpy(words, w.data(), sizeof(uint32_t) * std::min<size_t>(w.size(), (size_t)cap));
  });
}

int ksim_program_pack_blob(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                           const int32_t *f_selection, const int32_t *F_per_comp,
                           int32_t n_comp, int32_t num_f, const int32_t *dest_col,
                           int32_t num_outputs, uint8_t *out, int64_t cap, int64_t *n_bytes) {
  return guarded([&] {
    PackStats st;
    std::vector<uint32_t> w = pack_program(graphs, graphs_per_comp, f_selection, F_per_comp,
                                           n_comp, num_f, dest_col, num_outputs, st);
    double s8[8];
    stats_to8(st, w.size(), s8);
    copy_out(blob_from_words(w, s8), out, cap, n_bytes);
  });
}

int ksim_blob_check(const uint8_t *blob, int64_t n, double *stats8) {
  return guarded([&] {
    if (n < 0) throw std::invalid_argument("negative blob size");
    (void)words_from_blob(blob, (size_t)n, stats8);
  });
}

int ksim_program_serialize(const ksim_program *p, uint8_t *out, int64_t cap, int64_t *n_bytes) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("program is NULL");
    std::vector<uint32_t> w(p->n_words);
    hip_check(hipMemcpy(w.data(), p->d_prog, p->n_words * 4, hipMemcpyDeviceToHost),
              "download program");
    double s8[8];
    stats_to8(p->st, w.size(), s8);
    copy_out(blob_from_words(w, s8), out, cap, n_bytes);
  });
}

int ksim_program_create_from_blob(const uint8_t *blob, int64_t n, ksim_program **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (n < 0) throw std::invalid_argument("negative blob size");
    double s8[8];
    std::vector<uint32_t> w = words_from_blob(blob, (size_t)n, s8);
    *out = upload_program(w, stats_from8(s8, (int)w[PH_WMAX]));
  });
}

int ksim_program_info(const ksim_program *p, int64_t *info8) {
  return guarded([&] {
    if (!p || !info8) throw std::invalid_argument("NULL argument");
    const int64_t v[8] = {p->n_comp, p->num_f, p->num_outputs, p->tot_out,
                          p->wmax, (int64_t)p->n_words, p->device, p->out_words};
    std::memcpy(info8, v, sizeof(v));
  });
}

int ksim_tables(int32_t *tab, int64_t tab_cap, int32_t *pitab, int64_t pitab_cap,
                int32_t *sizes4) {
  return guarded([&] {
    if (tab)
      std::memcpy(tab, materialise_table(),
                  sizeof(int32_t) * std::min<int64_t>(tab_cap, 4 * TAB_DIGITS * TAB_N));
    if (pitab)
      std::memcpy(pitab, pi_table(), sizeof(int32_t) * std::min<int64_t>(pitab_cap, 4 * PI_MAX * PI_MAX));
    if (sizes4) {
      sizes4[0] = TAB_W; sizes4[1] = PI_MAX; sizes4[2] = TAB_DIGITS; sizes4[3] = TAB_DIGIT_BITS;
    }
  });
}

int ksim_device_info(int64_t *info4) {
  return guarded([&] {
    if (!info4) throw std::invalid_argument("NULL argument");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t pr;
    hip_check(hipGetDeviceProperties(&pr, dev), "hipGetDeviceProperties");
    info4[0] = pr.multiProcessorCount;
    info4[1] = (int64_t)pr.maxSharedMemoryPerMultiProcessor;
    info4[2] = pr.clockRate;
    info4[3] = (int64_t)pr.totalGlobalMem;
  });
}

}  // extern "C"

// ------------------------------------------------------------------ batch

namespace {

// The work list of a batch: per program its table row, and the items
// (program, component, shot range, term split) heaviest first.
struct BatchPlan {
  std::vector<BatchProg> progs;
  std::vector<BatchItem> items;
  int wmax = 1;            // parameter words per lane of the widest program
  int wl = 1;              // register words of the kernel (0: words in LDS)
  bool all_narrow = true;  // every program without 128-bit graphs
  bool approx = false;     // some program has approximate graphs (LDS-word kernel)
};

BatchPlan plan_batch(ksim_program *const *progs, int32_t n, const uint8_t *const *d_f,
                     const int32_t *f_cols, const int64_t *B, const uint64_t *seed,
                     uint8_t *const *d_out, int32_t wave_shots, double tile_ops) {
  if (n < 0) throw std::invalid_argument("n < 0");
  if (n > 0 && (!progs || !d_f || !f_cols || !B || !seed || !d_out))
    throw std::invalid_argument("NULL argument");
  if (wave_shots < 64 || wave_shots % 64) throw std::invalid_argument("wave_shots must be a multiple of 64");
  double total_ops = 0.0;
  for (int k = 0; k < n; ++k) {
    ksim_program *p = progs[k];
    if (!p) throw std::invalid_argument("NULL program in batch");
    if (B[k] < 0) throw std::invalid_argument("B < 0");
    if (B[k] > 0 && ((!d_f[k] && f_cols[k] > 0) || !d_out[k])) throw std::invalid_argument("NULL buffer");
    if (f_cols[k] < p->num_f) throw std::invalid_argument("f matrix has fewer columns than f_selection needs");
    total_ops += p->st.ops * (double)B[k];
  }
  BatchPlan plan;
  plan.progs.resize(n);
  struct Cost { double c; BatchItem it; };
  std::vector<Cost> items;
  // a tile should take each wave at most 1/8 of its share of the batch
  // (waves: 5 resident blocks of 4 waves per CU), so the heaviest-first
  // order leaves a short tail
  int dev = 0, cus = 256;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
  const double tile_target =
      tile_ops > 0 ? tile_ops : total_ops / (double)(cus * 5 * 4) / 8.0;
  for (int k = 0; k < n; ++k) {
    ksim_program *p = progs[k];
    plan.wmax = std::max(plan.wmax, p->wmax);
    plan.all_narrow = plan.all_narrow && p->all_narrow;
    plan.approx = plan.approx || p->approx;
    plan.progs[k] = BatchProg{p->d_prog, d_f[k], d_out[k], seed[k], B[k], f_cols[k],
                              p->num_outputs, p->tot_out, p->wl};
    for (int c = 0; c < p->n_comp; ++c) {
      const int L = split_for(p, c, tile_target);
      const double tc = tile_cost(p, c, L);
      // one item: wave_shots shots per group of L waves, 4 / L groups
      const int64_t span = (int64_t)wave_shots * (4 / L);
      for (int64_t s0 = 0; s0 < B[k]; s0 += span) {
        const int64_t ns = std::min<int64_t>(span, B[k] - s0);
        const int64_t rounds = ((ns + 63) / 64 + (4 / L) - 1) / (4 / L);
        items.push_back(Cost{tc * (double)rounds, BatchItem{k, c, s0, (int32_t)ns, L}});
      }
    }
  }
  plan.wl = plan.approx ? 0 : lane_words(plan.wmax);
  // heaviest items first: the dynamic cursor then balances the tail
  std::stable_sort(items.begin(), items.end(),
                   [](const Cost &x, const Cost &y) { return x.c > y.c; });
  if (items.size() >= 0xffffffffull) throw std::invalid_argument("too many work items");
  plan.items.resize(items.size());
  for (size_t i = 0; i < items.size(); ++i) plan.items[i] = items[i].it;
  return plan;
}

// One batched launch on `st`: clear the cursor and error word (a fill
// kernel, so the launch can be captured), then the persistent kernel.
void launch_batch(const BatchProg *d_progs, const BatchItem *d_items, uint32_t n_items,
                  uint32_t *d_cursor, int wl, int wmax, bool all_narrow, int blocks,
                  const DevTables &t, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_cursor, dim3(1), dim3(64), 0, st, d_cursor, d_cursor + 1);
  hip_check(hipGetLastError(), "k_reset_cursor");
  if (wl == 0) {
    const size_t dyn = lds_words_bytes(wmax);
    if (all_narrow) {
      allow_lds(reinterpret_cast<const void *>(&k_batch_lds<false>), dyn);
      hipLaunchKernelGGL((k_batch_lds<false>), dim3(blocks), dim3(kBlock), dyn, st, d_progs,
                         d_items, n_items, d_cursor, d_cursor + 1, t.view(), wmax);
    } else {
      allow_lds(reinterpret_cast<const void *>(&k_batch_lds<true>), dyn);
      hipLaunchKernelGGL((k_batch_lds<true>), dim3(blocks), dim3(kBlock), dyn, st, d_progs,
                         d_items, n_items, d_cursor, d_cursor + 1, t.view(), wmax);
    }
    hip_check(hipGetLastError(), "k_batch_lds");
    return;
  }
#define KSIM_BATCH_LAUNCH(W, WD)                                                               \
  hipLaunchKernelGGL((k_batch<W, WD>), dim3(blocks), dim3(kBlock), 0, st, d_progs, d_items,     \
                     n_items, d_cursor, d_cursor + 1, t.view())
#define KSIM_BATCH_CASES(W)                               \
  if (all_narrow) KSIM_BATCH_LAUNCH(W, false);            \
  else KSIM_BATCH_LAUNCH(W, true);
  switch (wl) {
    case 1: KSIM_BATCH_CASES(1); break;
    case 2: KSIM_BATCH_CASES(2); break;
    case 4: KSIM_BATCH_CASES(4); break;
    default: KSIM_BATCH_CASES(8); break;
  }
#undef KSIM_BATCH_CASES
#undef KSIM_BATCH_LAUNCH
  hip_check(hipGetLastError(), "k_batch");
}

}  // namespace

struct ksim_batch {
  std::vector<ksim_program *> progs;
  BatchProg *d_progs = nullptr;
  BatchItem *d_items = nullptr;
  uint32_t n_items = 0;
  uint32_t *d_cursor = nullptr;  // [0] cursor, [1] err
  int wave_shots = 64;
  int wmax = 1;  // parameter words per lane of the widest program of the batch
  int wl = 1;    // register words of its kernel (0: words in LDS)
  bool all_narrow = true;  // every program without 128-bit graphs
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_blocks = 0;
  hipStream_t graph_stream = nullptr;
};

extern "C" {

int ksim_program_split_hint(const ksim_program *p, int32_t comp, double tile_ops,
                            int32_t *split, double *tile_cost_ops) {
  return guarded([&] {
    if (!p || !split) throw std::invalid_argument("NULL argument");
    if (comp < 0 || comp >= p->n_comp) throw std::invalid_argument("component index out of range");
    *split = split_for(p, comp, tile_ops);
    if (tile_cost_ops) *tile_cost_ops = tile_cost(p, comp, *split);
  });
}


int ksim_batch_create(ksim_program *const *progs, int32_t n, const uint8_t *const *d_f,
                      const int32_t *f_cols, const int64_t *B, const uint64_t *seed,
                      uint8_t *const *d_out, int32_t wave_shots, double tile_ops,
                      ksim_batch **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("NULL argument");
    BatchPlan plan = plan_batch(progs, n, d_f, f_cols, B, seed, d_out, wave_shots, tile_ops);
    // owned until handed out, so every throw below releases it
    std::unique_ptr<ksim_batch, void (*)(ksim_batch *)> hold(new ksim_batch(), ksim_batch_destroy);
    ksim_batch *b = hold.get();
    b->wave_shots = wave_shots;
    b->progs.assign(progs, progs + n);
    b->wl = plan.wl;
    b->wmax = plan.wmax;
    b->all_narrow = plan.all_narrow;
    b->n_items = (uint32_t)plan.items.size();
    hip_check(hipMalloc(&b->d_progs, sizeof(BatchProg) * std::max(1, n)), "hipMalloc");
    hip_check(hipMalloc(&b->d_items, sizeof(BatchItem) * std::max<size_t>(1, plan.items.size())),
              "hipMalloc");
    hip_check(hipMalloc(&b->d_cursor, 64), "hipMalloc");
    if (n)
      hip_check(hipMemcpy(b->d_progs, plan.progs.data(), sizeof(BatchProg) * n,
                          hipMemcpyHostToDevice), "H2D");
    if (!plan.items.empty())
      hip_check(hipMemcpy(b->d_items, plan.items.data(), sizeof(BatchItem) * plan.items.size(),
                          hipMemcpyHostToDevice), "H2D");
    hip_check(hipMemset(b->d_cursor, 0, 64), "hipMemset");
    tables_for_current_device();
    *out = hold.release();
  });
}

static void batch_enqueue(ksim_batch *b, int blocks, hipStream_t st) {
  launch_batch(b->d_progs, b->d_items, b->n_items, b->d_cursor, b->wl, b->wmax, b->all_narrow,
               blocks, tables_for_current_device(), st);
}

int ksim_batch_run(ksim_batch *b, int32_t persistent_blocks, int32_t use_graph, void *stream) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("batch is NULL");
    if (persistent_blocks <= 0) throw std::invalid_argument("persistent_blocks must be > 0");
    hipStream_t st = (hipStream_t)stream;
    if (!use_graph) {
      batch_enqueue(b, persistent_blocks, st);
      return;
    }
    if (!b->exec || b->graph_blocks != persistent_blocks) {
      if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
      if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
      hipStream_t cs;
      hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "BeginCapture");
      batch_enqueue(b, persistent_blocks, cs);
      hip_check(hipStreamEndCapture(cs, &b->graph), "EndCapture");
      hip_check(hipStreamDestroy(cs), "hipStreamDestroy");
      hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "GraphInstantiate");
      b->graph_blocks = persistent_blocks;
    }
    hip_check(hipGraphLaunch(b->exec, st), "hipGraphLaunch");
  });
}

int ksim_batch_check(ksim_batch *b) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("batch is NULL");
    uint32_t e = 0;
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hip_check(hipMemcpy(&e, b->d_cursor + 1, 4, hipMemcpyDeviceToHost), "read err");
    if (e) throw RangeError(err_text(e));
  });
}

void ksim_batch_destroy(ksim_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_progs) (void)hipFree(b->d_progs);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_cursor) (void)hipFree(b->d_cursor);
  delete b;
}

}  // extern "C"

// ------------------------------------------------------- batch, no allocation
// ksim_batch_sample_static keeps the work list, the program table, the work
// cursor and the error word of its launch in `scratch`, a fixed array in
// device memory (one copy per GPU, held until the process ends), so a call
// allocates nothing.  Layout: program table, work items, cursor words, each
// on a 256-byte boundary.
namespace {

constexpr size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kStaticProgsOff = 0;
constexpr size_t kStaticItemsOff = round256(sizeof(BatchProg) * KSIM_STATIC_MAX_PROGS);
constexpr size_t kStaticCursorOff =
    kStaticItemsOff + round256(sizeof(BatchItem) * (size_t)KSIM_STATIC_MAX_ITEMS);
constexpr size_t kStaticBytes = kStaticCursorOff + 256;

}  // namespace

// the array itself; host code uses only the per-device address that
// hipGetSymbolAddress returns for it (its host-side name is no device address)
__device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch;  // device -> address of its `scratch`

// the address of this device's `scratch`, looked up once per device with the
// device current
unsigned char *scratch_for_current_device() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  void *ptr = nullptr;
  hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!ptr) throw DeviceError("hipGetSymbolAddress(scratch) returned NULL");
  g_scratch[dev] = (unsigned char *)ptr;
  return (unsigned char *)ptr;
}

// the current device's lookup tables, without building them: they exist once
// a program has been created on the device
DevTables tables_if_built() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_tab_mu);
  auto it = g_tabs.find(dev);
  if (it == g_tabs.end())
    throw std::invalid_argument("no program has been created on this device yet");
  return it->second;
}

}  // namespace

extern "C" {

int ksim_batch_sample_static(ksim_program *const *progs, int32_t n, const uint8_t *const *d_f,
                             const int32_t *f_cols, const int64_t *B, const uint64_t *seed,
                             uint8_t *const *d_out, int32_t wave_shots, double tile_ops,
                             int32_t persistent_blocks, void *stream) {
  return guarded([&] {
    if (n > KSIM_STATIC_MAX_PROGS)
      throw std::invalid_argument("more than KSIM_STATIC_MAX_PROGS (" +
                                  std::to_string(KSIM_STATIC_MAX_PROGS) +
                                  ") programs in one call without allocation");
    if (persistent_blocks <= 0) throw std::invalid_argument("persistent_blocks must be > 0");
    BatchPlan plan = plan_batch(progs, n, d_f, f_cols, B, seed, d_out, wave_shots, tile_ops);
    if (plan.items.size() > (size_t)KSIM_STATIC_MAX_ITEMS)
      throw std::invalid_argument(
          std::to_string(plan.items.size()) + " work items exceed KSIM_STATIC_MAX_ITEMS (" +
          std::to_string(KSIM_STATIC_MAX_ITEMS) + "); use larger wave_shots or fewer shots");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    for (int k = 0; k < n; ++k)
      if (progs[k]->device != dev)
        throw std::invalid_argument("a program lives on another device than the current one");
    if (plan.items.empty()) return;
    const DevTables t = tables_if_built();
    unsigned char *base = scratch_for_current_device();
    BatchProg *dp = (BatchProg *)(base + kStaticProgsOff);
    BatchItem *di = (BatchItem *)(base + kStaticItemsOff);
    uint32_t *dc = (uint32_t *)(base + kStaticCursorOff);
    hipStream_t st = (hipStream_t)stream;
    hip_check(hipMemcpyAsync(dp, plan.progs.data(), sizeof(BatchProg) * (size_t)n,
                             hipMemcpyHostToDevice, st), "H2D programs");
    hip_check(hipMemcpyAsync(di, plan.items.data(), sizeof(BatchItem) * plan.items.size(),
                             hipMemcpyHostToDevice, st), "H2D items");
    launch_batch(dp, di, (uint32_t)plan.items.size(), dc, plan.wl, plan.wmax, plan.all_narrow,
                 persistent_blocks, t, st);
    // the next call rewrites the same array: finish this launch first
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    uint32_t e = 0;
    hip_check(hipMemcpy(&e, dc + 1, sizeof(e), hipMemcpyDeviceToHost), "read err");
    if (e) throw RangeError(err_text(e));
  });
}

}  // extern "C"

// ------------------------------------------------------------------ scheduler
// A persistent-kernel launcher over an arbitrary list of (program slot,
// component, shot range) items per launch: the device side of the serving
// loop.  Slots bind a resident program to its resident f matrix, output and
// seed.  Launches rotate over KSIM_SCHED_BUFS staging sets (pinned host
// buffers, device buffers and a work cursor each), so up to that many
// launches can be in flight on different streams and overlap on the device:
// the blocks of launch k+1 start as those of launch k drain.
constexpr int KSIM_SCHED_BUFS = 4;

struct ksim_sched {
  int max_slots = 0, max_items = 0;
  std::vector<BatchProg> slots;
  std::vector<int> slot_wmax, slot_ncomp;  // parameter words per lane, components
  std::vector<char> slot_approx;            // program with approximate graphs
  std::vector<char> slot_narrow;
  // per slot and component: per-wave cost of a 64-shot tile at split 1, 2, 4
  std::vector<std::vector<std::array<double, 3>>> slot_cost;
  BatchProg *h_progs[KSIM_SCHED_BUFS] = {};
  BatchItem *h_items[KSIM_SCHED_BUFS] = {};
  BatchProg *d_progs[KSIM_SCHED_BUFS] = {};
  BatchItem *d_items[KSIM_SCHED_BUFS] = {};
  hipEvent_t done[KSIM_SCHED_BUFS] = {};  // recorded after the launch's kernel
  bool done_valid[KSIM_SCHED_BUFS] = {};
  uint32_t *d_cursor = nullptr;  // [2k] cursor of set k, [2k+1] scratch; [64] error word
  int next = 0;
};

extern "C" {

int ksim_sched_create(int32_t max_slots, int32_t max_items, ksim_sched **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    if (max_slots <= 0 || max_items <= 0) throw std::invalid_argument("sizes must be > 0");
    auto *q = new ksim_sched();
    q->max_slots = max_slots;
    q->max_items = max_items;
    q->slots.assign(max_slots, BatchProg{nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 1});
    q->slot_wmax.assign(max_slots, 0);
    q->slot_approx.assign(max_slots, 0);
    q->slot_ncomp.assign(max_slots, 0);
    q->slot_narrow.assign(max_slots, 1);
    q->slot_cost.assign(max_slots, {});
    try {
      for (int k = 0; k < KSIM_SCHED_BUFS; ++k) {
        hip_check(hipHostMalloc((void **)&q->h_progs[k], sizeof(BatchProg) * max_slots), "hipHostMalloc");
        hip_check(hipHostMalloc((void **)&q->h_items[k], sizeof(BatchItem) * max_items), "hipHostMalloc");
        hip_check(hipMalloc(&q->d_progs[k], sizeof(BatchProg) * max_slots), "hipMalloc");
        hip_check(hipMalloc(&q->d_items[k], sizeof(BatchItem) * max_items), "hipMalloc");
        hip_check(hipEventCreateWithFlags(&q->done[k], hipEventDisableTiming), "hipEventCreate");
      }
      hip_check(hipMalloc(&q->d_cursor, 65 * sizeof(uint32_t)), "hipMalloc");
      hip_check(hipMemset(q->d_cursor, 0, 65 * sizeof(uint32_t)), "hipMemset");
      tables_for_current_device();
    } catch (...) {
      ksim_sched_destroy(q);
      throw;
    }
    *out = q;
  });
}

int ksim_sched_set_slot(ksim_sched *q, int32_t slot, ksim_program *p, const uint8_t *d_f,
                        int32_t f_cols, int64_t B, uint64_t seed, uint8_t *d_out) {
  return guarded([&] {
    if (!q || !p) throw std::invalid_argument("NULL argument");
    if (slot < 0 || slot >= q->max_slots) throw std::invalid_argument("slot out of range");
    if (B < 0) throw std::invalid_argument("B < 0");
    // work items address a request's shots with 32-bit offsets
    if (B > INT32_MAX) throw std::invalid_argument("more than 2^31 - 1 shots in one request");
    if (f_cols < p->num_f) throw std::invalid_argument("f matrix has fewer columns than f_selection needs");
    if (B > 0 && ((!d_f && f_cols > 0) || !d_out)) throw std::invalid_argument("NULL buffer");
    q->slots[slot] = BatchProg{p->d_prog, d_f, d_out, seed, B, f_cols, p->num_outputs,
                               p->tot_out, p->wl};
    q->slot_wmax[slot] = p->wmax;
    q->slot_approx[slot] = p->approx ? 1 : 0;
    q->slot_ncomp[slot] = p->n_comp;
    q->slot_narrow[slot] = p->all_narrow ? 1 : 0;
    q->slot_cost[slot].assign(p->n_comp, {0.0, 0.0, 0.0});
    for (int c = 0; c < p->n_comp; ++c)
      for (int j = 0; j < 3; ++j) q->slot_cost[slot][c][j] = tile_cost(p, c, 1 << j);
  });
}

int ksim_sched_clear_slot(ksim_sched *q, int32_t slot) {
  return guarded([&] {
    if (!q) throw std::invalid_argument("NULL argument");
    if (slot < 0 || slot >= q->max_slots) throw std::invalid_argument("slot out of range");
    q->slots[slot] = BatchProg{nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 1};
    q->slot_wmax[slot] = 0;
    q->slot_approx[slot] = 0;
    q->slot_ncomp[slot] = 0;
    q->slot_narrow[slot] = 1;
    q->slot_cost[slot].clear();
  });
}

// items: n rows of (slot, component, first shot, shot count, term split),
// int32 each.
int ksim_sched_launch(ksim_sched *q, const int32_t *items, int32_t n, int32_t blocks,
                      void *stream) {
  return guarded([&] {
    if (!q) throw std::invalid_argument("NULL argument");
    if (n < 0 || n > q->max_items) throw std::invalid_argument("item count out of range");
    if (blocks <= 0) throw std::invalid_argument("blocks must be > 0");
    if (n == 0) return;
    if (!items) throw std::invalid_argument("items is NULL");
    hipStream_t st = (hipStream_t)stream;
    const int k = q->next;
    q->next = (q->next + 1) % KSIM_SCHED_BUFS;
    // the set's previous launch (possibly on another stream) must be over
    if (q->done_valid[k]) hip_check(hipEventSynchronize(q->done[k]), "hipEventSynchronize");
    int wmax = 1;
    bool narrow = true, approx = false;
    for (int i = 0; i < n; ++i) {
      const int32_t slot = items[5 * i], comp = items[5 * i + 1];
      const int32_t shot0 = items[5 * i + 2], count = items[5 * i + 3];
      const int32_t split = items[5 * i + 4];
      if (split != 1 && split != 2 && split != 4)
        throw std::invalid_argument("item term split must be 1, 2 or 4");
      if (slot < 0 || slot >= q->max_slots || !q->slots[slot].prog)
        throw std::invalid_argument("item names an empty slot");
      const BatchProg &bp = q->slots[slot];
      if (comp < 0 || comp >= q->slot_ncomp[slot])
        throw std::invalid_argument("item component index out of range");
      if (shot0 < 0 || count <= 0 || (int64_t)shot0 + count > bp.B)
        throw std::invalid_argument("item shot range outside its request");
      wmax = std::max(wmax, q->slot_wmax[slot]);
      approx = approx || q->slot_approx[slot];
      narrow = narrow && q->slot_narrow[slot];
      q->h_items[k][i] = BatchItem{slot, comp, shot0, count, split};
    }
    // heaviest items first (per-wave tile cost x rounds), as in a batch: the
    // cursor then hands the long tiles out while the short ones fill the end
    {
      std::vector<std::pair<double, int>> order(n);
      for (int i = 0; i < n; ++i) {
        const BatchItem &it = q->h_items[k][i];
        const int G = 4 / it.split;
        const double rounds = (double)(((it.count + 63) / 64 + G - 1) / G);
        const int j = it.split == 1 ? 0 : it.split == 2 ? 1 : 2;
        order[i] = {q->slot_cost[it.prog][it.comp][j] * rounds, i};
      }
      std::stable_sort(order.begin(), order.end(),
                       [](const std::pair<double, int> &x, const std::pair<double, int> &y) {
                         return x.first > y.first;
                       });
      std::vector<BatchItem> tmp(q->h_items[k], q->h_items[k] + n);
      for (int i = 0; i < n; ++i) q->h_items[k][i] = tmp[order[i].second];
    }
    std::memcpy(q->h_progs[k], q->slots.data(), sizeof(BatchProg) * q->max_slots);
    hip_check(hipMemcpyAsync(q->d_progs[k], q->h_progs[k], sizeof(BatchProg) * q->max_slots,
                             hipMemcpyHostToDevice, st), "H2D slots");
    hip_check(hipMemcpyAsync(q->d_items[k], q->h_items[k], sizeof(BatchItem) * n,
                             hipMemcpyHostToDevice, st), "H2D items");
    DevTables t = tables_for_current_device();
    uint32_t *cursor = q->d_cursor + 2 * k;
    const int wl = approx ? 0 : lane_words(wmax);
    if (wl == 0) {
      // a component wider than the register words, or a program with
      // approximate graphs: every item's parameter words in LDS
      hipLaunchKernelGGL(k_reset_cursor, dim3(1), dim3(64), 0, st, cursor, cursor + 1);
      hip_check(hipGetLastError(), "k_reset_cursor");
      const size_t dyn = lds_words_bytes(wmax);
      if (narrow) {
        allow_lds(reinterpret_cast<const void *>(&k_batch_lds<false>), dyn);
        hipLaunchKernelGGL((k_batch_lds<false>), dim3(blocks), dim3(kBlock), dyn, st,
                           q->d_progs[k], q->d_items[k], (uint32_t)n, cursor, q->d_cursor + 64,
                           t.view(), wmax);
      } else {
        allow_lds(reinterpret_cast<const void *>(&k_batch_lds<true>), dyn);
        hipLaunchKernelGGL((k_batch_lds<true>), dim3(blocks), dim3(kBlock), dyn, st,
                           q->d_progs[k], q->d_items[k], (uint32_t)n, cursor, q->d_cursor + 64,
                           t.view(), wmax);
      }
      hip_check(hipGetLastError(), "k_batch_lds");
    } else {
      hipLaunchKernelGGL(k_reset_cursor, dim3(1), dim3(64), 0, st, cursor, cursor + 1);
      hip_check(hipGetLastError(), "k_reset_cursor");
#define KSIM_SCHED_LAUNCH(W, WD)                                                               \
  hipLaunchKernelGGL((k_batch<W, WD>), dim3(blocks), dim3(kBlock), 0, st, q->d_progs[k],        \
                     q->d_items[k], (uint32_t)n, cursor, q->d_cursor + 64, t.view())
#define KSIM_SCHED_CASES(W)                        \
  if (narrow) KSIM_SCHED_LAUNCH(W, false);         \
  else KSIM_SCHED_LAUNCH(W, true);
      switch (wl) {
        case 1: KSIM_SCHED_CASES(1); break;
        case 2: KSIM_SCHED_CASES(2); break;
        case 4: KSIM_SCHED_CASES(4); break;
        default: KSIM_SCHED_CASES(8); break;
      }
#undef KSIM_SCHED_CASES
#undef KSIM_SCHED_LAUNCH
      hip_check(hipGetLastError(), "k_batch");
    }
    hip_check(hipEventRecord(q->done[k], st), "hipEventRecord");
    q->done_valid[k] = true;
  });
}

int ksim_sched_check(ksim_sched *q, void *stream) {
  return guarded([&] {
    if (!q) throw std::invalid_argument("NULL argument");
    hipStream_t st = (hipStream_t)stream;
    uint32_t e = 0;
    for (int k = 0; k < KSIM_SCHED_BUFS; ++k)
      if (q->done_valid[k]) hip_check(hipEventSynchronize(q->done[k]), "hipEventSynchronize");
    hip_check(hipMemcpyAsync(&e, q->d_cursor + 64, 4, hipMemcpyDeviceToHost, st), "read err");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (e) throw RangeError(err_text(e));
  });
}

void ksim_sched_destroy(ksim_sched *q) {
  if (!q) return;
  for (int k = 0; k < KSIM_SCHED_BUFS; ++k) {
    if (q->done[k]) {
      if (q->done_valid[k]) (void)hipEventSynchronize(q->done[k]);
      (void)hipEventDestroy(q->done[k]);
    }
    if (q->h_progs[k]) (void)hipHostFree(q->h_progs[k]);
    if (q->h_items[k]) (void)hipHostFree(q->h_items[k]);
    if (q->d_progs[k]) (void)hipFree(q->d_progs[k]);
    if (q->d_items[k]) (void)hipFree(q->d_items[k]);
  }
  if (q->d_cursor) (void)hipFree(q->d_cursor);
  delete q;
}

}  // extern "C"

// ------------------------------------------------------------------ noise
struct ksim_noise {
  int device = -1;  // device holding the tables (uploaded on first device use)
  NoiseHost h;
  uint64_t *d_skip = nullptr;
  NoiseChan *d_chan = nullptr;
  uint64_t *d_cdf = nullptr;
  uint32_t *d_pat = nullptr;
  uint8_t *d_rows = nullptr;  // staging for the host-pointer API
  size_t rows_cap = 0;

  NoiseDev dev(uint64_t seed) const {
    NoiseDev d;
    d.skip = d_skip; d.chan = d_chan; d.cdf = d_cdf; d.pat = d_pat;
    d.seed = seed; d.C = h.C; d.wf = h.wf;
    return d;
  }
};

namespace {

template <class T>
T *upload_vec(const std::vector<T> &v, const char *what) {
  T *d = nullptr;
  const size_t n = std::max<size_t>(v.size(), 1) * sizeof(T);
  hip_check(hipMalloc(&d, n), what);
  if (!v.empty()) hip_check(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), what);
  return d;
}

// upload the tables to the current device on first use (a noise model is
// built on the host, so it can be inspected without a device)
void noise_on_device(ksim_noise *nz) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (nz->device == dev) return;
  if (nz->device >= 0) throw std::invalid_argument("noise model used on a second device");
  nz->d_skip = upload_vec(nz->h.skip, "upload noise skip table");
  nz->d_chan = upload_vec(nz->h.chan, "upload noise channels");
  nz->d_cdf = upload_vec(nz->h.cdf, "upload noise cdf");
  nz->d_pat = upload_vec(nz->h.pat, "upload noise patterns");
  nz->device = dev;
}

void launch_noise_rows(ksim_noise *nz, int64_t B, int64_t shot_base, uint64_t seed,
                       uint8_t *d_out, int32_t f_cols, hipStream_t st) {
  if (f_cols < nz->h.num_f) throw std::invalid_argument("f_cols is smaller than the noise model's f bits");
  if (B <= 0 || nz->h.num_f == 0) return;
  noise_on_device(nz);
  const int64_t nb = (B + kBlock - 1) / kBlock;
  if (nb > 0x7fffffff) throw std::invalid_argument("too many shots for one launch");
  hipLaunchKernelGGL(k_noise_rows, dim3((unsigned)nb), dim3(kBlock), 0, st, nz->dev(seed), B,
                     shot_base, nz->h.num_f, f_cols, d_out);
  hip_check(hipGetLastError(), "k_noise_rows launch");
}

void sample_noisy_device(ksim_program *p, ksim_noise *nz, int64_t B, int64_t shot_base,
                         uint64_t seed, uint64_t noise_seed, const double *d_U, uint8_t *d_out,
                         hipStream_t st) {
  if (!p || !nz) throw std::invalid_argument("NULL argument");
  if (B < 0) throw std::invalid_argument("B < 0");
  if (nz->h.num_f < p->num_f)
    throw std::invalid_argument("the noise model has fewer f bits than the program selects");
  if (B > 0 && !d_out) throw std::invalid_argument("NULL buffer");
  reset_err_async(p, st);
  noise_on_device(nz);
  const NoiseDev nd = nz->dev(noise_seed);
  if (nd.C == 0) {
    // nothing can fire: every f bit is 0
    const int32_t fc = std::max(p->num_f, 1);
    ensure(p->d_f, p->f_cap, (size_t)B * fc);
    fill_u8_async(p->d_f, 0, B * fc, st);
    launch_sample(p, p->d_f, B, fc, shot_base, seed, d_U, d_out, st);
  } else {
    launch_sample(p, nullptr, B, 0, shot_base, seed, d_U, d_out, st, 0, &nd);
  }
}

}  // namespace

extern "C" {

int ksim_noise_create(const double *probs, const int32_t *n_outcomes, const uint8_t *patterns,
                      int32_t n_channels, int32_t num_f, ksim_noise **out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("out is NULL");
    auto *nz = new ksim_noise();
    try {
      nz->h = build_noise(probs, n_outcomes, patterns, n_channels, num_f);
    } catch (...) {
      ksim_noise_destroy(nz);
      throw;
    }
    *out = nz;
  });
}

int ksim_noise_info(const ksim_noise *nz, int64_t *info4) {
  return guarded([&] {
    if (!nz || !info4) throw std::invalid_argument("NULL argument");
    info4[0] = nz->h.C;
    info4[1] = nz->h.num_f;
    info4[2] = nz->h.wf;
    info4[3] = (int64_t)nz->h.cdf.size();
  });
}

int ksim_noise_tables(const ksim_noise *nz, uint64_t *skip, uint64_t *acc, uint32_t *n_out,
                      uint32_t *off_cdf, uint32_t *off_pat, uint64_t *cdf, uint32_t *pat,
                      int32_t *source) {
  return guarded([&] {
    if (!nz) throw std::invalid_argument("noise is NULL");
    const NoiseHost &h = nz->h;
    if (skip) std::memcpy(skip, h.skip.data(), h.skip.size() * 8);
    for (int c = 0; c < h.C; ++c) {
      if (acc) acc[c] = h.chan[c].acc;
      if (n_out) n_out[c] = h.chan[c].n_out;
      if (off_cdf) off_cdf[c] = h.chan[c].off_cdf;
      if (off_pat) off_pat[c] = h.chan[c].off_pat;
      if (source) source[c] = h.source[c];
    }
    if (cdf && !h.cdf.empty()) std::memcpy(cdf, h.cdf.data(), h.cdf.size() * 8);
    if (pat && !h.pat.empty()) std::memcpy(pat, h.pat.data(), h.pat.size() * 4);
  });
}

int ksim_noise_sample_device(ksim_noise *nz, int64_t B, int64_t shot_base, uint64_t seed,
                             uint8_t *d_out, int32_t f_cols, void *stream) {
  return guarded([&] {
    if (!nz) throw std::invalid_argument("noise is NULL");
    if (B < 0) throw std::invalid_argument("B < 0");
    if (B > 0 && nz->h.num_f > 0 && !d_out) throw std::invalid_argument("NULL buffer");
    launch_noise_rows(nz, B, shot_base, seed, d_out, f_cols, (hipStream_t)stream);
  });
}

int ksim_noise_sample(ksim_noise *nz, int64_t B, int64_t shot_base, uint64_t seed,
                      uint8_t *out) {
  return guarded([&] {
    if (!nz) throw std::invalid_argument("noise is NULL");
    if (B < 0) throw std::invalid_argument("B < 0");
    const size_t n = (size_t)B * nz->h.num_f;
    if (n == 0) return;
    if (!out) throw std::invalid_argument("NULL buffer");
    if (n > nz->rows_cap) {
      if (nz->d_rows) hip_check(hipFree(nz->d_rows), "hipFree");
      nz->d_rows = nullptr;
      nz->rows_cap = 0;
      hip_check(hipMalloc(&nz->d_rows, n), "hipMalloc(noise rows)");
      nz->rows_cap = n;
    }
    launch_noise_rows(nz, B, shot_base, seed, nz->d_rows, nz->h.num_f, nullptr);
    hip_check(hipMemcpy(out, nz->d_rows, n, hipMemcpyDeviceToHost), "D2H noise rows");
  });
}

int ksim_program_sample_noisy_device(ksim_program *p, ksim_noise *nz, int64_t B,
                                     int64_t shot_base, uint64_t seed, uint64_t noise_seed,
                                     const double *d_U, uint8_t *d_out, void *stream) {
  return guarded([&] {
    sample_noisy_device(p, nz, B, shot_base, seed, noise_seed, d_U, d_out, (hipStream_t)stream);
  });
}

int ksim_program_sample_noisy(ksim_program *p, ksim_noise *nz, int64_t B, uint64_t seed,
                              uint64_t noise_seed, uint8_t *out) {
  return guarded([&] {
    if (!p || !nz) throw std::invalid_argument("NULL argument");
    if (B < 0) throw std::invalid_argument("B < 0");
    if (B == 0) return;
    if (!out) throw std::invalid_argument("NULL buffer");
    const size_t ob = (size_t)B * p->num_outputs;
    ensure(p->d_out, p->out_cap, ob);
    if (p->tot_out < p->num_outputs) fill_u8_async(p->d_out, 0, (int64_t)ob, nullptr);
    sample_noisy_device(p, nz, B, 0, seed, noise_seed, nullptr, p->d_out, nullptr);
    hip_check(hipMemcpy(out, p->d_out, ob, hipMemcpyDeviceToHost), "D2H out");
    check_err_sync(p, nullptr);
  });
}

void ksim_noise_destroy(ksim_noise *nz) {
  if (!nz) return;
  if (nz->d_skip) (void)hipFree(nz->d_skip);
  if (nz->d_chan) (void)hipFree(nz->d_chan);
  if (nz->d_cdf) (void)hipFree(nz->d_cdf);
  if (nz->d_pat) (void)hipFree(nz->d_pat);
  if (nz->d_rows) (void)hipFree(nz->d_rows);
  delete nz;
}

}  // extern "C"
