/*
 * ksim_hip.h — C ABI of the MI355X (gfx950) FlatProgram sampler.
 *
 * Drop-in native layer for the reference's `kokkos_sim` module
 * (src/cpp/ksim/kokkos_sim.cpp:244-252, bound by nanobind):
 *
 *   reference (nanobind)                      this ABI
 *   ----------------------------------------  -------------------------------------
 *   Component(graphs: list[dict], num_f)      ksim_component_create
 *     kokkos_sim.cpp:67-151
 *   Component.sample(f_sel, seed)             ksim_component_sample
 *     kokkos_sim.cpp:188-240
 *   Component.eval_abs(graph, params)         ksim_component_eval_abs
 *     kokkos_sim.cpp:157-186
 *   Component.n_out                           ksim_component_n_out
 *     kokkos_sim.cpp:250-251
 *   (Python GC of Component)                  ksim_program_destroy
 *
 * plus whole-program entry points that replace the per-component Python loop
 * of KokkosProgramSampler.sample (src/python/ksim/sample/sampler.py:57-67)
 * with one launch, a fed-uniforms variant matching sample_flat's draw order
 * (src/python/ksim/sample/evaluate.py:73-82), an exact evaluator for parity
 * checks, device-pointer variants for resident inputs, and a batched
 * persistent launcher over many programs.
 *
 * Conventions: every function returns 0 on success and a negative error code
 * otherwise; ksim_last_error() returns the message of the last failure on the
 * calling thread.  Host pointers are borrowed for the duration of the call
 * only (the library copies what it keeps).  Calls are synchronous unless they
 * take a `stream` argument.  Parameter and mask bytes are read by their
 * lowest bit.
 */
#ifndef KSIM_HIP_H
#define KSIM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KSIM_OK 0
#define KSIM_E_INVALID -1     /* bad argument (reference: std::invalid_argument) */
#define KSIM_E_UNSUPPORTED -2 /* program outside the exact range / features   */
#define KSIM_E_DEVICE -3      /* HIP runtime error                            */
#define KSIM_E_RANGE -4       /* a device-side exact-range check failed        */

/* One FlatScalarGraph, host memory, C-contiguous; the 19 arrays of the
 * reference's _graph_to_dict (src/python/ksim/sample/sampler.py:21-43).
 * Shapes: G terms, P params, T_* padded family widths. */
typedef struct ksim_graph_view {
  int32_t G, P, T_np, T_hp, T_pp, T_ph;
  const uint8_t *np_phases;      /* (G, T_np) */
  const uint8_t *np_params;      /* (G, T_np, P) */
  const int32_t *np_counts;      /* (G) */
  const uint8_t *hp_coeffs;      /* (G, T_hp) */
  const uint8_t *hp_params;      /* (G, T_hp, P) */
  const uint8_t *pp_psi_c;       /* (G, T_pp) */
  const uint8_t *pp_psi_p;       /* (G, T_pp, P) */
  const uint8_t *pp_phi_c;       /* (G, T_pp) */
  const uint8_t *pp_phi_p;       /* (G, T_pp, P) */
  const uint8_t *ph_alpha;       /* (G, T_ph) */
  const uint8_t *ph_alpha_p;     /* (G, T_ph, P) */
  const uint8_t *ph_beta;        /* (G, T_ph) */
  const uint8_t *ph_beta_p;      /* (G, T_ph, P) */
  const int32_t *ph_counts;      /* (G) */
  const uint8_t *pf_phase_idx;   /* (G) */
  const int32_t *pf_floatfactor; /* (G, 4) */
  const int32_t *pf_power2;      /* (G) */
  const float *pf_approx_f32;    /* (G, 2) complex64 as float pairs */
  int32_t pf_has_approx;
} ksim_graph_view;

typedef struct ksim_program ksim_program;

/* ---- kokkos_sim.Component equivalents --------------------------------- */

/* Build and upload one component from its n_graphs = n_out + 1 graphs.
 * Replaces the Component constructor (kokkos_sim.cpp:67-151). */
int ksim_component_create(const ksim_graph_view *graphs, int32_t n_graphs,
                          int32_t num_f, ksim_program **out);

/* Sample B shots: f_sel (B, num_f) u8 row-major -> out (B, max(n_out,1)) u8.
 * Same seeding and uniform stream as kokkos_sim.cpp:214-229 (xorshift64 per
 * shot).  num_f mismatch is checked by the caller binding (as the reference
 * does at kokkos_sim.cpp:193-194); here f_cols must equal num_f. */
int ksim_component_sample(ksim_program *p, const uint8_t *f_sel, int64_t B,
                          int32_t f_cols, uint64_t seed, uint8_t *out);

/* |amplitude| of graph `graph` for explicit param rows (B, P) u8 -> out (B)
 * float64.  Replaces Component.eval_abs (kokkos_sim.cpp:157-186). */
int ksim_component_eval_abs(ksim_program *p, int32_t graph,
                            const uint8_t *params, int64_t B, int32_t P,
                            double *out);

/* n_out of component 0 (kokkos_sim.cpp:250-251). */
int32_t ksim_component_n_out(const ksim_program *p);

/* ---- whole-program entry points --------------------------------------- */

/* Build and upload every component of a FlatProgram.
 *   graphs            all graphs, component-major (sum(graphs_per_comp))
 *   graphs_per_comp   (n_comp) n_out_c + 1
 *   f_selection       concatenated f_selection arrays (sum(F_per_comp))
 *   F_per_comp        (n_comp)
 *   num_f             columns of the f_params matrix the program needs at
 *                     least (1 + the largest f_selection entry); calls pass
 *                     the actual column count f_cols >= num_f
 *   dest_col          (sum n_out_c) output column of each component output,
 *                     i.e. the inverse of argsort(output_order)
 *                     (src/python/ksim/sample/sampler.py:65-67)
 *   num_outputs       columns of the result */
int ksim_program_create(const ksim_graph_view *graphs,
                        const int32_t *graphs_per_comp,
                        const int32_t *f_selection, const int32_t *F_per_comp,
                        int32_t n_comp, int32_t num_f, const int32_t *dest_col,
                        int32_t num_outputs, ksim_program **out);

/* f_params (B, f_cols) u8 -> out (B, num_outputs) u8; f_cols >= num_f.  Component k draws its
 * uniforms from seed + 0x9E37*k, as KokkosProgramSampler.sample does
 * (sampler.py:63). */
int ksim_program_sample(ksim_program *p, const uint8_t *f_params, int64_t B,
                        int32_t f_cols, uint64_t seed, uint8_t *out);

/* As ksim_program_sample but with fed uniforms U (B, sum n_out_c) float64,
 * columns component-major then output-major: the order sample_flat draws
 * rng.random(B) in (evaluate.py:73-82). */
int ksim_program_sample_uniforms(ksim_program *p, const uint8_t *f_params,
                                 int64_t B, int32_t f_cols, const double *U,
                                 uint8_t *out);

/* Exact value of graph `graph` of component `comp` at param rows (B, P) u8:
 * coeffs (B, 4, 2) int64 as (lo, hi) words of four signed 128-bit integers in
 * the basis (1, w, i, conj w), exp2 (B): value = coeffs * 2^exp2.  Not
 * canonicalised.  No reference counterpart: the parity-check entry point. */
int ksim_program_eval_exact(ksim_program *p, int32_t comp, int32_t graph,
                            const uint8_t *params, int64_t B, int32_t P,
                            int64_t *coeffs, int32_t *exp2);

/* Device-pointer sampling for inputs already resident in HBM.  d_f (B, f_cols)
 * u8, d_U optional (NULL: seeded stream), d_out (B, num_outputs) u8.
 * shot_base offsets the shot index that keys the per-shot stream, so a range
 * split of one request samples exactly what a single call would.  Enqueued
 * on `stream` (hipStream_t, NULL = default); returns without waiting.  A
 * 16-byte aligned d_f is read once per 64-row tile, coalesced; any other
 * alignment is accepted and read row by row. */
int ksim_program_sample_device(ksim_program *p, const uint8_t *d_f, int64_t B,
                               int32_t f_cols, int64_t shot_base, uint64_t seed,
                               const double *d_U, uint8_t *d_out, void *stream);

/* Packed output epilogue: rows of ceil(num_outputs / 32) little-endian u32
 * words, bit c of a row = output column c (after the output_order
 * permutation), i.e. numpy's packbits(bitorder="little") of the u8 row padded
 * to whole words — 8x less D2H than the u8 rows of sampler.py:65-67.  Same
 * shots and stream as ksim_program_sample / _sample_device.  The device
 * variant zeroes d_out itself (on `stream`) when rows are assembled from
 * several components. */
int ksim_program_sample_packed(ksim_program *p, const uint8_t *f_params, int64_t B,
                               int32_t f_cols, uint64_t seed, uint32_t *out);
int ksim_program_sample_device_packed(ksim_program *p, const uint8_t *d_f, int64_t B,
                                      int32_t f_cols, int64_t shot_base, uint64_t seed,
                                      const double *d_U, uint32_t *d_out, void *stream);

/* Launch shape of the plain kernel: threads per block (256) and term
 * split, the waves that share one 64-shot tile's term sums: 1 (each lane
 * evaluates every term of its own shot), 2 or 4, or 0 (default: chosen per
 * launch, splitting launches whose tiles cannot fill the GPU's resident
 * wave slots, or whose waves would leave a nearly empty last round of
 * slots).  The waves of a split tile meet in LDS once per
 * graph of several terms.  No reference counterpart (the reference runs
 * one thread per shot). */
int ksim_program_set_tiles(ksim_program *p, int32_t block_threads,
                           int32_t term_split);

/* Error word of the last launch (0 = none), for callers of the async API. */
int ksim_program_check(ksim_program *p, void *stream);

/* 1 when every graph of the program is exact, 0 when some graph carries
 * approximate (complex64) floatfactors (pf_has_approx,
 * src/cpp/ksim/kokkos_sim.cpp:123-129): such a graph's terms are exact in
 * Z[w] but are cast to float64, multiplied by their factors and summed in
 * float64 in term order, as the reference does (zx_eval.hpp:168-178), and
 * ksim_program_eval_exact rejects it.  -1 for NULL. */
int32_t ksim_program_exact(const ksim_program *p);

/* Bytes of device memory the program holds. */
int64_t ksim_program_device_bytes(const ksim_program *p);

/* Stats of the packed program: [n_graphs, n_terms, n_records, words,
 * ops_per_shot (algorithmic int64 ops per shot, all components),
 * largest |n'| (the (1+sqrt2) exponent) a term can reach, terms whose
 * |n'| may exceed the single-digit table rows, bits of the exact sum the
 * widest graph may need (at most 128)]. */
int ksim_program_stats(const ksim_program *p, double *stats8);

void ksim_program_destroy(ksim_program *p);

/* ---- wire format: a packed program as one validated blob ---------------- */

/* Program blob layout, version 6 (little-endian; version 3, 4 and 5 blobs
 * are read too): "KSIMPRG\0", u32 version,
 * u32 header bytes (128), u64 word count, u64 FNV-1a 64 of the words,
 * 8 f64 stats (ksim_program_stats order), u32 n_comp / num_f / num_outputs /
 * total outputs / lane words, 12 zero bytes, then the u32 program words.
 * Replaces the reference's per-process packing of the graph arrays
 * (src/cpp/ksim/kokkos_sim.cpp:91-151, src/python/ksim/sample/sampler.py:21-43)
 * with a format that uploads without re-packing.  Each writer stores up to
 * `cap` bytes to `out` and the full size to *n_bytes (out may be NULL). */
int ksim_program_serialize(const ksim_program *p, uint8_t *out, int64_t cap,
                           int64_t *n_bytes);

/* Host-only: pack a FlatProgram (arguments as ksim_program_create) straight
 * into a blob. */
int ksim_program_pack_blob(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                           const int32_t *f_selection, const int32_t *F_per_comp,
                           int32_t n_comp, int32_t num_f, const int32_t *dest_col,
                           int32_t num_outputs, uint8_t *out, int64_t cap,
                           int64_t *n_bytes);

/* Validate a blob (magic, version, sizes, checksum, and every offset, count
 * and record size the kernels follow); stats8 (optional) receives its stats.
 * Host-only. */
int ksim_blob_check(const uint8_t *blob, int64_t n, double *stats8);

/* Validate and upload a blob as a program (no packing). */
int ksim_program_create_from_blob(const uint8_t *blob, int64_t n, ksim_program **out);

/* [n_comp, num_f, num_outputs, total outputs, lane words, program words,
 * device, packed output row words] of a program. */
int ksim_program_info(const ksim_program *p, int64_t *info8);

/* ---- device-side noise: the f bits drawn on the GPU ---------------------- */

typedef struct ksim_noise ksim_noise;

/* A noise model: n_channels channels, channel c with n_outcomes[c] (a power
 * of two) outcome probabilities in `probs` (concatenated, outcome 0 = no
 * error) and, for outcomes 1.., the f bits each flips as rows of num_f bytes
 * in `patterns` (concatenated over channels, n_outcomes[c] - 1 rows each).
 * Replaces ChannelSampler (src/python/ksim/compile/channels.py:503-658):
 * same per-channel fire probability and conditional outcome distribution,
 * drawn per shot on the device from a SplitMix64 stream keyed by
 * (seed, shot index) instead of numpy's PCG64 on the host.  The tables are
 * built on the host here and uploaded on first device use. */
int ksim_noise_create(const double *probs, const int32_t *n_outcomes, const uint8_t *patterns,
                      int32_t n_channels, int32_t num_f, ksim_noise **out);

/* [channels kept, num_f, pattern words per row, cdf entries]. */
int ksim_noise_info(const ksim_noise *nz, int64_t *info4);

/* Copy the sampling tables (sizes from ksim_noise_info; any pointer may be
 * NULL): skip (C+1 u64), per kept channel acc (u64), n_out, off_cdf,
 * off_pat (u32) and its input index `source` (i32), cdf (u64), pat (u32). */
int ksim_noise_tables(const ksim_noise *nz, uint64_t *skip, uint64_t *acc, uint32_t *n_out,
                      uint32_t *off_cdf, uint32_t *off_pat, uint64_t *cdf, uint32_t *pat,
                      int32_t *source);

/* Draw f rows for shots [shot_base, shot_base + B): d_out (B, f_cols) u8,
 * f_cols >= num_f, columns [0, num_f) written. */
int ksim_noise_sample_device(ksim_noise *nz, int64_t B, int64_t shot_base, uint64_t seed,
                             uint8_t *d_out, int32_t f_cols, void *stream);

/* Host variant: out (B, num_f) u8, the rows of ksim_noise_sample_device. */
int ksim_noise_sample(ksim_noise *nz, int64_t B, int64_t shot_base, uint64_t seed,
                      uint8_t *out);

/* Sample a program with its f bits drawn inside the kernel (no f matrix in
 * HBM, no host noise): the same result as ksim_program_sample_device on the
 * rows ksim_noise_sample_device(nz, B, shot_base, noise_seed) produces.
 * The noise model must cover the program's f bits. */
int ksim_program_sample_noisy_device(ksim_program *p, ksim_noise *nz, int64_t B,
                                     int64_t shot_base, uint64_t seed, uint64_t noise_seed,
                                     const double *d_U, uint8_t *d_out, void *stream);
int ksim_program_sample_noisy(ksim_program *p, ksim_noise *nz, int64_t B, uint64_t seed,
                              uint64_t noise_seed, uint8_t *out);

void ksim_noise_destroy(ksim_noise *nz);

/* ---- batched persistent path ------------------------------------------ */

typedef struct ksim_batch ksim_batch;

/* A batch of n programs sampled by one persistent kernel.  Each program k
 * gets its own resident f matrix d_f[k] (B[k], f_cols[k]) u8 and output
 * d_out[k] (B[k], num_outputs_k) u8 and seed[k].  Work items are (program,
 * component, shot range, term split): a persistent 256-thread block takes
 * one item at a time; with term split L its four waves form 4/L groups
 * whose L waves share each graph's terms on one 64-shot tile.  wave_shots
 * (a multiple of 64) is the shots one group covers per item.  The split
 * of each component is ksim_program_split_hint's for a tile budget of
 * tile_ops algorithmic ops per wave (<= 0: 1/8 of one wave's share of the
 * batch, with 5 blocks per CU; a huge budget turns the split off). */
int ksim_batch_create(ksim_program *const *progs, int32_t n,
                      const uint8_t *const *d_f, const int32_t *f_cols,
                      const int64_t *B,
                      const uint64_t *seed, uint8_t *const *d_out,
                      int32_t wave_shots, double tile_ops, ksim_batch **out);

/* Run one batched launch on `stream`.  use_graph != 0 captures the launch
 * (work-list reset + persistent kernel) into a hipGraph on first use and
 * replays it afterwards. */
int ksim_batch_run(ksim_batch *b, int32_t persistent_blocks, int32_t use_graph,
                   void *stream);

int ksim_batch_check(ksim_batch *b);

/* Term split for component `comp` of a program in the persistent kernels:
 * the smallest L in {1, 2, 4} whose 64-shot tile costs each of its L waves
 * at most `tile_ops` algorithmic ops while keeping at least 60% of their
 * work useful (graphs with one term are evaluated by every wave of the
 * group).  tile_cost_ops (optional) receives that per-wave tile cost.  No
 * reference counterpart (the reference runs one thread per shot). */
int ksim_program_split_hint(const ksim_program *p, int32_t comp, double tile_ops,
                            int32_t *split, double *tile_cost_ops);
void ksim_batch_destroy(ksim_batch *b);

/* ---- batch without allocation ------------------------------------------- */

/* Sizes of the fixed device array ksim_batch_sample_static works in. */
#define KSIM_STATIC_MAX_PROGS 256
#define KSIM_STATIC_MAX_ITEMS 65536

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  One batched persistent launch over n programs, the
 * same work items and results as ksim_batch_create + ksim_batch_run with
 * the same arguments (the caller passes every input and output as for
 * ksim_batch_create), but the work list, program table, work cursor and
 * error word live in a fixed array in device memory, one per GPU (about
 * 1.6 MB, held until the process ends), so the call allocates nothing.
 * Limits: n <= KSIM_STATIC_MAX_PROGS and at most KSIM_STATIC_MAX_ITEMS work
 * items (a component of B shots at term split L takes
 * ceil(B / (wave_shots * 4 / L)) items); larger calls are rejected.  The
 * programs must live on the current device, where creating them has built
 * the lookup tables.  Blocks until the launch has finished and reports
 * device range errors.  Every call on one GPU shares its array: calls that
 * overlap on one GPU (from several threads or streams) race. */
int ksim_batch_sample_static(ksim_program *const *progs, int32_t n,
                             const uint8_t *const *d_f, const int32_t *f_cols,
                             const int64_t *B, const uint64_t *seed,
                             uint8_t *const *d_out, int32_t wave_shots,
                             double tile_ops, int32_t persistent_blocks,
                             void *stream);

/* ---- scheduler: the device side of the serving loop ---------------------- */

typedef struct ksim_sched ksim_sched;

/* A persistent-kernel launcher whose work list is given per launch.  No
 * reference counterpart (the reference has no serving path). */
int ksim_sched_create(int32_t max_slots, int32_t max_items, ksim_sched **out);

/* Bind slot `slot` to a resident program with its resident f matrix
 * d_f (B, f_cols) u8, output d_out (B, num_outputs) u8 and seed; a request's
 * shots use the same per-shot streams as ksim_program_sample_device with
 * shot_base 0, however they are split into items.  B < 2^31 (items carry
 * 32-bit shot offsets). */
int ksim_sched_set_slot(ksim_sched *q, int32_t slot, ksim_program *p,
                        const uint8_t *d_f, int32_t f_cols, int64_t B,
                        uint64_t seed, uint8_t *d_out);
int ksim_sched_clear_slot(ksim_sched *q, int32_t slot);

/* Enqueue one persistent launch on `stream` over n items, each 5 int32:
 * (slot, component, first shot, shot count, term split L in {1, 2, 4}).
 * A persistent block takes one item: 4/L groups of L waves, group g
 * sampling the 64-shot tiles g, g + 4/L, ... of the item (so an item of
 * 256/L shots gives every wave one tile).  Returns after staging.
 * Launches rotate over four staging sets (own buffers and work cursor), so
 * up to four launches may be in flight on different streams; launch k+4
 * waits for launch k to finish before reusing its set. */
int ksim_sched_launch(ksim_sched *q, const int32_t *items, int32_t n,
                      int32_t blocks, void *stream);

/* Synchronise `stream` and report a device exact-range error, if any. */
int ksim_sched_check(ksim_sched *q, void *stream);
void ksim_sched_destroy(ksim_sched *q);

/* ---- host-only inspection (no device needed) ---------------------------- */

/* Pack a program exactly as ksim_program_create does, without uploading:
 * writes up to `cap` words to `words`, the full word count to *n_words and,
 * when stats8 is not NULL, the ksim_program_stats figures.
 * No reference counterpart: lets the packer be checked on a CPU-only host. */
int ksim_program_pack(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                      const int32_t *f_selection, const int32_t *F_per_comp,
                      int32_t n_comp, int32_t num_f, const int32_t *dest_col,
                      int32_t num_outputs, uint32_t *words, int64_t cap,
                      int64_t *n_words, double *stats8);

/* Copy the kernels' lookup tables.  tab: D planes of (2W+1)*4*8 entries of
 * 4 int32, entry ((n+W)*4+r)*8+s of plane d holding digit d (base 2^G,
 * balanced) of the coefficients of w^s (1+w)^r (1+sqrt2)^n; pitab:
 * PI_MAX*PI_MAX entries of 4 int32 (pi^a pibar^b).  sizes4 receives
 * W, PI_MAX, D and G.  No reference counterpart: lets the packed format be
 * interpreted on a CPU-only host. */
int ksim_tables(int32_t *tab, int64_t tab_cap, int32_t *pitab, int64_t pitab_cap,
                int32_t *sizes4);

/* ---- misc --------------------------------------------------------------- */

const char *ksim_last_error(void);

/* Device properties used by the tuner: [CUs, LDS bytes/CU, clock kHz,
 * HBM bytes]. */
int ksim_device_info(int64_t *info4);

#ifdef __cplusplus
}
#endif

#endif /* KSIM_HIP_H */
