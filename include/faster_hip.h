/*
 * libfaster_hip -- C ABI of the MI355X (gfx950) linkage hot path.
 *
 * Conventions
 *   - Every entry point returns an int status (FH_OK == 0). On failure a
 *     thread-local message is available from fh_last_error(). No C++
 *     exception crosses this boundary.
 *   - Pointers documented as "device" point to HIP device memory owned by the
 *     caller (the Python host hands in torch tensors' data_ptr()); the library
 *     never frees caller memory. Internal scratch comes from HIP's
 *     stream-ordered pool on the caller's stream and is returned before the
 *     call ends.
 *   - Work is ordered on `stream` (a hipStream_t, e.g. torch's current
 *     stream). Entry points that report a size to the host (n_hits, n_out,
 *     ...) synchronise the stream once the size is known; the work that
 *     fills the output may still be running when they return (later work on
 *     `stream` is ordered after it).
 *   - Output arenas have a capacity. When the result does not fit, the call
 *     still returns FH_OK, reports the required size, and writes nothing
 *     beyond the capacity; the caller grows the arena and calls again.
 *   - Pair-list keys ("level keys") are uint64 values id * 4 + level, where id
 *     is the reference's pair id (row_A * n_B + row_B for linkage,
 *     row * n + col with row > col for deduplication) and level is 1 or 2.
 *     Lists are ascending unless stated otherwise.
 *
 * Each declaration names the reference code (fast-er 0.3.1, paths relative to
 * the reference repository root) whose role it takes over.
 */
#ifndef FASTER_HIP_H
#define FASTER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* fh_stream_t; /* hipStream_t */

enum {
  FH_OK = 0,
  FH_ERR_HIP = 1,         /* a HIP runtime call failed */
  FH_ERR_ARG = 2,         /* invalid argument */
  FH_ERR_UNSUPPORTED = 3, /* input outside what this build supports */
  FH_ERR_INTERNAL = 4
};

enum { FH_MODE_LINK = 0, FH_MODE_DEDUP = 1 };

/* Message of the last failed call on this thread ("" if none). */
const char* fh_last_error(void);
/* ABI version (major * 10000 + minor * 100 + patch). */
int fh_version(void);

/* Measurement helper, no counterpart in the reference: launches `blocks` x 256
 * threads that run `iters` rounds of the comparison kernel's match recurrence
 * (three VALU ops per dword on register operands, 8 independent dwords per
 * thread) on `stream`, and reports the lane-ops launched. `out` holds
 * blocks * 256 uint32 (one per thread, written so no work is dead). The
 * caller times the launch; bench.py --valu-probe does, to check the VALU
 * peak its roofline uses. */
int fh_probe_valu(uint32_t* out, int32_t blocks, int32_t iters, double* lane_ops, fh_stream_t stream);

/* Work accounting of one fh_jw_levels call (host struct, optional). */
typedef struct {
  int64_t pairs_logical;   /* value pairs the call covers */
  int64_t pairs_kernel;    /* pairs evaluated by the comparison kernel (not pruned by length) */
  int64_t pairs_pruned;    /* pairs skipped because their lengths cannot reach level 1 */
  int64_t pairs_generic;   /* pairs evaluated by the generic kernels (empty or > 64 bytes; any length) */
  double alg_lane_ops;     /* algorithmic VALU lane-ops of the kernel pairs */
  int32_t alphabet;        /* byte values present on both sides */
  int32_t launches;        /* kernel launches of the comparison stage */
  double kernel_ms;        /* device time of the comparison kernels (HIP events) */
  int32_t method;          /* FH_JW_METHOD_* used */
  int32_t reserved;
  int64_t pairs_candidates; /* pairs handed to the scoring kernel (greedy count passed) */
  double pass_ms;          /* device time of the all-pairs pass kernels alone (HIP events) */
  double eval_ms;          /* device time of the candidate scoring kernels (HIP events) */
  double frozen_model_ops; /* sum over the evaluated pairs of 8 * len1 + 24: a fixed work model of a
                              non-packed recurrence (8 ops per string-1 character + 24 per pair) */
  double prefix_skipped_ops; /* FH_JW_METHOD_PREFIX: lane-ops of alg_lane_ops that the reuse of shared
                                prefixes skips (recurrence steps not recomputed); the executed
                                algorithmic work is alg_lane_ops - prefix_skipped_ops */
} fh_jw_stats;

/*
 * Comparison algorithm of fh_jw_levels (process-wide; both give identical
 * results):
 *   FH_JW_METHOD_TILES: the bit-parallel greedy match of every pair (VALU),
 *     one string-1 value per lane.
 *   FH_JW_METHOD_PREFIX (default): the same bit-parallel match with the
 *     string-1 values walked one after another by each wave, whose lanes
 *     cover the string-2 tile; the match state of a prefix shared with the
 *     previous (sorted) value is reused. String-2 values longer than 32
 *     bytes (and values over 64 bytes) use the TILES kernel, and so does the
 *     whole call when more than about 55 byte values occur on both sides
 *     (the prefix form's 256-dword tile then exceeds the LDS); the stats'
 *     `method` reports the form that ran.
 * No counterpart in the reference (one CUDA kernel scores every pair,
 * faster/comparison.py:133-186); the switch exists to time and check the
 * forms against each other.
 */
#define FH_JW_METHOD_TILES 0
#define FH_JW_METHOD_PREFIX 2
int fh_jw_set_method(int method);
int fh_jw_get_method(void);

/*
 * All-pairs Jaro-Winkler levels over two sets of byte strings.
 *
 * Replaces the comparison kernel faster/comparison.py:133-186 (score of every
 * pair, per-pair work arrays in global memory) together with the
 * thresholding/compaction at faster/comparison.py:321-338; in
 * FH_MODE_DEDUP it replaces faster/deduplication.py:132-196 and :513-529.
 *
 * Strings: chars_* (device bytes) with offsets_* (device int64, n + 1
 * entries). Level 1: lower_thr <= s < upper_thr, level 2: s >= upper_thr, with
 * s the reference's fp32 score (p is the prefix weight) and comparisons in
 * fp32.
 *   FH_MODE_LINK : every pair (a, b), string 1 = A value. Key
 *                  (a * n_b + b) * 4 + level.
 *   FH_MODE_DEDUP: set A only (chars_b/offsets_b/n_b ignored); every
 *                  unordered pair of distinct indices x != y, string 1 = the
 *                  lexicographically larger value. Key (x * n_a + y) * 4 +
 *                  level with x the index of string 1. Equal indices (the
 *                  reference's forced diagonal) are not emitted.
 * hits: device uint64[capacity], unordered. *n_hits (host) receives the number
 * of hits (possibly > capacity, see the conventions above).
 */
int fh_jw_levels(const uint8_t* chars_a, const int64_t* offsets_a, int64_t n_a,
                 const uint8_t* chars_b, const int64_t* offsets_b, int64_t n_b, int mode,
                 float p, float lower_thr, float upper_thr, uint64_t* hits, int64_t capacity,
                 int64_t* n_hits, fh_jw_stats* stats, fh_stream_t stream);

/*
 * Batched blocked linkage (one persistent, graph-captured launch).
 *
 * Replaces the reference's blocking recipe, one Comparison.fit per block with
 * the Counts summed (docs/source/usage.rst:34-76), by a single plan over all
 * blocks. Rows of both sides carry a block id (device int32 block_a[n_a],
 * block_b[n_b], values in [0, n_blocks)); pairs are compared only within a
 * block. Fuzzy variable v has device byte strings chars_a[v]/offsets_a[v]
 * (n_a strings) and chars_b[v]/offsets_b[v] (n_b); exact variable q has
 * device int64 value ids ids_a[q]/ids_b[q] (equal ids = equal values). The
 * host arrays of device pointers are read during the call only.
 *
 * fh_blocked_plan_create sorts rows by (block, length), builds waves, tiles
 * and largest-first work items, and allocates every buffer (it synchronises).
 * fh_blocked_run enqueues one run on the plan's stream, ordered after
 * `stream` and before later work on `stream`; with use_graph the run is
 * captured into a hipGraph on first use and replayed afterwards.
 * fh_blocked_counts waits for the last run and copies out the per-block
 * histogram hist[block * n_patterns + code] (host int64) of every pair with at
 * least one fuzzy level > 0 (pattern code over the fuzzy then the exact
 * variables, first variable slowest), the number of such pairs, an overflow
 * flag (the pair table was too small: rerun with more hash_slots) and the
 * device time of the run in ms. Patterns whose fuzzy digits are all 0 are not
 * counted here (their sizes follow from value-group counts per block).
 * Strings may have any length: values of 1..64 bytes go through the
 * bit-parallel passes, pairs with a longer value through a scoring kernel of
 * their own (flags in registers up to 1,024 bytes, in plan-allocated scratch
 * beyond). Empty strings pair at level 0; the thresholds must be above 0.
 */
typedef struct {
  int32_t tile_width;     /* largest string-2 tile width in dwords: 32, 64, 128 or 256; the plan
                             halves it while the tile rows of the largest alphabet and the
                             survivor queues exceed 64 KB of LDS */
  int32_t tiles_per_item; /* tiles per work item */
  int32_t waves_per_item; /* string-1 waves per work item (1..8) */
  int32_t wgs_per_cu;     /* persistent workgroups per compute unit */
  int64_t hash_slots;     /* pair-table slots (0: 2 * (n_a + n_b), at least 2^22) */
} fh_blocked_tuning;

int fh_blocked_plan_create(int n_fuzzy, const uint8_t* const* chars_a, const int64_t* const* offsets_a,
                           int64_t n_a, const uint8_t* const* chars_b, const int64_t* const* offsets_b,
                           int64_t n_b, const int32_t* block_a, const int32_t* block_b, int32_t n_blocks,
                           int n_exact, const int64_t* const* ids_a, const int64_t* const* ids_b, float p,
                           float lower_thr, float upper_thr, const fh_blocked_tuning* tuning,
                           void** plan, fh_stream_t stream);
int fh_blocked_run(void* plan, int use_graph, fh_stream_t stream);
int fh_blocked_counts(void* plan, int64_t* hist, int64_t* n_union, int* overflow, double* ms);
/*
 * The plain per-block launch loop over the same plan, for comparison with the
 * batched launch: the reference's recipe runs one Comparison.fit per block
 * (docs/source/usage.rst:41-76); here block after block gets one table build
 * per fuzzy variable over its tiles and one comparison launch per string-length
 * part over its work items, enqueued on the plan's stream with no graph and no
 * host synchronisation. The counts accumulate on the device and are read by
 * fh_blocked_counts exactly as after fh_blocked_run (same results).
 */
int fh_blocked_run_per_block(void* plan, fh_stream_t stream);
/* info (host double[8]): work items (short, long strings), patterns, table
 * slots, pairs evaluated, algorithmic lane-ops, grid size, LDS bytes. */
int fh_blocked_info(void* plan, double* info);
int fh_blocked_plan_destroy(void* plan);

/*
 * Groups the n rows of a string column by value on the device.
 *
 * Replaces the host np.unique of faster/comparison.py:377,395 and
 * faster/deduplication.py:420 with the stable argsort of the inverse at
 * faster/comparison.py:380-391. chars/offsets (device, offsets int64[n + 1])
 * hold the rows' UTF-8 bytes. Outputs (device, caller-sized): uchars
 * (capacity offsets[n] bytes) and uoffsets (int64[n + 1]) the unique values
 * in byte-lexicographic order (shorter prefix first: np.unique's order for
 * str); starts (int64[n + 1]) and rows (int64[n]) list the rows of unique
 * value u as rows[starts[u] .. starts[u + 1]), ascending. *n_unique and
 * *unique_bytes are host outputs. Synchronizes the stream.
 */
int fh_group_strings(const uint8_t* chars, const int64_t* offsets, int64_t n, uint8_t* uchars,
                     int64_t* uoffsets, int64_t* starts, int64_t* rows, int64_t* n_unique,
                     int64_t* unique_bytes, fh_stream_t stream);

/*
 * fh_group_strings for k string columns (1..16) in one device sort: column
 * c has n[c] rows in chars[c]/offsets[c] and its outputs in uchars[c],
 * uoffsets[c], starts[c], rows[c] (sized as for fh_group_strings), with the
 * same contents as k separate calls; n_unique[c] and unique_bytes[c] (host).
 * The rows of all columns are sorted together by (column, value), so the
 * columns share the sort's launches and host round trips (two).
 * Synchronizes the stream.
 */
int fh_group_strings_multi(int k, const uint8_t* const* chars, const int64_t* const* offsets, const int64_t* n,
                           uint8_t* const* uchars, int64_t* const* uoffsets, int64_t* const* starts,
                           int64_t* const* rows, int64_t* n_unique, int64_t* unique_bytes, fh_stream_t stream);

/*
 * Groups the n rows of a numeric column by value on the device.
 *
 * Replaces the host np.unique(..., return_inverse=True) of an exact column
 * (faster/comparison.py:519,535,548; faster/deduplication.py:645) with the
 * stable argsort of its inverse. values (device, n entries) hold int64 values
 * (kind FH_VALUES_I64; narrower integers and booleans widened by the caller)
 * or the bit patterns of float64 values (FH_VALUES_F64; every NaN is one
 * value and -0.0 equals +0.0, as in np.unique). Outputs (device):
 * group[n] the value id of every row (ids in ascending value order),
 * starts[n_groups + 1] and rows[n] the rows of value g as rows[starts[g] ..
 * starts[g + 1]), ascending. *n_groups (host). Synchronises the stream.
 */
#define FH_VALUES_I64 0
#define FH_VALUES_F64 1
int fh_group_values(const int64_t* values, int64_t n, int kind, int64_t* group, int64_t* starts,
                    int64_t* rows, int64_t* n_groups, fh_stream_t stream);

/*
 * The B side of a grouping of the concatenated rows [A rows; B rows] (as
 * fh_group_values or fh_group_strings give it, rows ascending within a
 * group): starts_b[n_groups + 1] and rows_b[n_b] list the B rows (numbered
 * from 0) of group g as rows_b[starts_b[g] .. starts_b[g + 1]), ascending;
 * with group non-null (n_a + n_b entries) the group of every row is written
 * too. Replaces the joint unique and the count gathers of
 * faster/comparison.py:548-580. Device outputs; no synchronisation.
 */
int fh_groups_side_b(const int64_t* starts, const int64_t* rows, int64_t n_groups, int64_t n_a, int64_t n_b,
                     int64_t* starts_b, int64_t* rows_b, int64_t* group, fh_stream_t stream);

/*
 * Unique-value pair keys -> original-row pair keys (linkage).
 *
 * Replaces faster/comparison.py:429-453 (count gather, cumsum) and the
 * expansion kernel faster/comparison.py:196-245. hits: keys
 * (ua * n_ub + ub) * 4 + level. Unique value u of side A owns rows
 * rows_a[starts_a[u] .. starts_a[u + 1]) (device int64). Output keys
 * (row_a * n_b + row_b) * 4 + level, unordered; *n_out (host) = total.
 */
int fh_expand_link(const uint64_t* hits, int64_t n_hits, int64_t n_ub, const int64_t* starts_a,
                   const int64_t* rows_a, const int64_t* starts_b, const int64_t* rows_b,
                   int64_t n_b, uint64_t* out, int64_t capacity, int64_t* n_out,
                   fh_stream_t stream);

/*
 * The same expansion with the output keys ascending (the expansion plus the
 * level-list sort, faster/comparison.py:429-458 and the cp.sort of the
 * levels' indices). n_ua unique values on side A, n_a A rows (rows not listed
 * in rows_a pair with nothing). When A values repeat (at least twice as many
 * record pairs as (value, B row) pairs), each value's (B row, level) list is
 * built and sorted once and copied per row in row order, instead of sorting
 * every record pair. capacity protocol as fh_expand_link; out must not alias
 * hits.
 */
int fh_expand_link_sorted(const uint64_t* hits, int64_t n_hits, int64_t n_ua, int64_t n_ub,
                          const int64_t* starts_a, const int64_t* rows_a, int64_t n_a,
                          const int64_t* starts_b, const int64_t* rows_b, int64_t n_b, uint64_t* out,
                          int64_t capacity, int64_t* n_out, fh_stream_t stream);

/*
 * Unique-value pair keys -> original-row pair keys (deduplication).
 *
 * Replaces the count and expansion kernels faster/deduplication.py:206-243
 * and :253-338 (with :550-580). hits: keys (x * n_u + y) * 4 + level over
 * unique values; a pair of distinct values expands to every row pair, keyed
 * max(r1, r2) * n + min(r1, r2). When diag_level > 0, every unique value
 * with two or more rows also contributes its within-group pairs at that level
 * (the reference's diagonal, whose score is forced to 1). Output unordered.
 */
int fh_expand_dedup(const uint64_t* hits, int64_t n_hits, int64_t n_u, const int64_t* starts,
                    const int64_t* rows, int64_t n, int diag_level, uint64_t* out,
                    int64_t capacity, int64_t* n_out, fh_stream_t stream);

/*
 * Exact-equality pairs (linkage), ascending.
 *
 * Replaces faster/comparison.py:500-600 (joint unique, the count == 2 test,
 * the expansion kernel and the sort). group_a[i] (device int64) is the value
 * id of A row i, or -1 when no B row holds the value; B rows holding value g
 * are rows_b[starts_b[g] .. starts_b[g + 1]) in ascending order. Output keys
 * (i * n_b + j) * 4 + 1, ascending.
 */
int fh_exact_link(const int64_t* group_a, int64_t n_a, const int64_t* starts_b,
                  const int64_t* rows_b, int64_t n_groups, int64_t n_b, uint64_t* out,
                  int64_t capacity, int64_t* n_out, fh_stream_t stream);

/*
 * Exact-equality pairs (deduplication), ascending.
 *
 * Replaces faster/deduplication.py:628-701 (output_count/repeat/triangular
 * decode kernel :345-391 and the sort). group[i] is the value id of row i,
 * or negative for a row without a value (it pairs with nothing); rows of
 * value g are rows[starts[g] .. starts[g + 1]) ascending (starts[n_groups]
 * rows in all). Output keys
 * (i * n + j) * 4 + 1 for every j < i with an equal value, ascending.
 */
int fh_exact_dedup(const int64_t* group, int64_t n, const int64_t* starts, const int64_t* rows,
                   int64_t n_groups, uint64_t* out, int64_t capacity, int64_t* n_out,
                   fh_stream_t stream);

/*
 * Numeric-distance levels. No counterpart in the reference, which compares
 * only by Jaro-Winkler or equality (faster/comparison.py:346-600); these give
 * the numeric comparisons of a pattern with three levels. Values are float64
 * (device); NaN is a missing value and pairs with nothing. Two values x, y
 * are within distance d when max(x, y) <= fl(min(x, y) + d) (float64, round
 * to nearest): level 2 within d2, level 1 within d1 but not d2
 * (0 <= d2 <= d1). Output keys (id * 4 + level), unordered, into an arena as
 * above; *n_out (host) = total. Synchronises the stream.
 *   fh_numeric_link : every pair (i, j) of A x B, id i * n_b + j.
 *   fh_numeric_dedup: every pair i != j of one column, id
 *                     max(i, j) * n + min(i, j).
 */
int fh_numeric_link(const double* values_a, int64_t n_a, const double* values_b, int64_t n_b, double d2,
                    double d1, uint64_t* out, int64_t capacity, int64_t* n_out, fh_stream_t stream);
int fh_numeric_dedup(const double* values, int64_t n, double d2, double d1, uint64_t* out, int64_t capacity,
                     int64_t* n_out, fh_stream_t stream);

/*
 * Missing levels of a pattern-major pair list. No counterpart in the
 * reference (it has no missing values; faster/estimation.py:70 fixes the
 * levels). ids (device int64[n]) and pattern_offsets (device
 * int64[n_patterns + 1]) are fh_pattern_merge's output over variables whose
 * last level is "missing" (never listed); miss_a / miss_b (device uint32) hold
 * per row a bit per variable, set when the row's value is missing. A pair's
 * code gets add[v] (host int64[n_vars]: the missing level times the
 * variable's pattern stride) for every variable missing on either of its rows
 * (rows id / id_div and id % id_div; ids <= max_id). ids_out (device
 * int64[n]) receives the pairs pattern-major again, each pattern ascending;
 * offsets_out (host int64[n_patterns + 1]) its boundaries. Synchronises.
 */
int fh_pattern_add_missing(const int64_t* ids, const int64_t* pattern_offsets, int n_patterns, int64_t n,
                           const uint32_t* miss_a, const uint32_t* miss_b, int64_t id_div, int64_t max_id,
                           int n_vars, const int64_t* add, int64_t* ids_out, int64_t* offsets_out,
                           fh_stream_t stream);

/*
 * Ascending sort of uint64 keys on bits [begin_bit, end_bit).
 *
 * Replaces cp.sort at faster/comparison.py:456,487,588 and
 * faster/deduplication.py:577,614,689.
 */
int fh_sort_u64(uint64_t* keys, int64_t n, int begin_bit, int end_bit, fh_stream_t stream);

/*
 * The same sort from `in` into `out` (distinct buffers of n keys); `in` is
 * overwritten. With an odd number of 8-bit passes (key widths of 1-8, 17-24,
 * 33-40 bits ...) nothing is copied, where the in-place form ends with a copy.
 *
 * Replaces the same cp.sort calls as fh_sort_u64 (faster/comparison.py:456,
 * 487,588; faster/deduplication.py:577,614,689), whose unsorted input the
 * reference drops after sorting.
 */
int fh_sort_u64_into(uint64_t* in, uint64_t* out, int64_t n, int begin_bit, int end_bit, fh_stream_t stream);

/*
 * Stable split of a sorted level-key list into per-level id lists.
 *
 * The reference keeps one list per level (faster/comparison.py:498). ids
 * (device int64[n]) receives the ids of level 1 first, then level 2, ...,
 * each ascending; level_offsets (host int64[n_levels + 1]) the boundaries,
 * level 0 first (always empty).
 */
int fh_split_levels(const uint64_t* keys, int64_t n, int n_levels, int64_t* ids,
                    int64_t* level_offsets, fh_stream_t stream);

/*
 * Membership mask of a sorted list in another sorted list.
 *
 * Replaces the binary search kernel faster/search.py:8-50 (inside = 1: the
 * mask intersect() compacts with, inside = 0: the setdiff() mask).
 */
int fh_sorted_member_mask(const int64_t* arr1, int64_t n1, const int64_t* arr2, int64_t n2,
                          int inside, uint8_t* mask, fh_stream_t stream);

/*
 * Pattern merge: per-variable level keys -> per-pattern id lists.
 *
 * Replaces the setdiff/intersect cascade faster/comparison.py:695-725
 * (identical at faster/deduplication.py:773-803, with faster/search.py:57-157)
 * and the list sizes behind Counts (faster/comparison.py:732-748).
 * var_keys: host array of n_vars device pointers to ascending level-key lists
 * (var_sizes host int64); var_levels: levels per variable (level 0 included).
 * The pattern code of a pair is sum_v level_v * stride_v with the first
 * variable slowest (itertools.product order). ids (device int64[capacity])
 * receives the pairs of pattern 1, then pattern 2, ..., each ascending;
 * pattern_offsets (host int64[n_patterns + 1]) the boundaries (pattern 0, the
 * all-zero pattern, is empty). *n_union (host) = number of distinct pairs.
 */
int fh_pattern_merge(int n_vars, const uint64_t* const* var_keys, const int64_t* var_sizes,
                     const int32_t* var_levels, int64_t* ids, int64_t capacity,
                     int64_t* pattern_offsets, int64_t* n_union, fh_stream_t stream);

/*
 * Pattern merge with exact variables given as per-row value ids.
 *
 * As fh_pattern_merge for the n_vars listed variables, followed in pattern
 * order by n_exact two-level exact variables that are not materialised as
 * pair lists: the level of pair id is 1 when ids_a[k][id / id_div] ==
 * ids_b[k][id % id_div] (device int64 arrays; for deduplication ids_b[k] =
 * ids_a[k] and id_div = n). Replaces the materialised exact lists of
 * faster/deduplication.py:628-701 (and faster/comparison.py:500-600) inside
 * the merge when those lists would be too large to build (the 2.6e10 pairs of
 * a 2M-row deduplication on birth year). Only pairs listed by some listed
 * variable are emitted; patterns whose listed digits are all 0 stay empty and
 * their sizes follow from value-group counts on the host.
 */
int fh_pattern_merge_rowids(int n_vars, const uint64_t* const* var_keys, const int64_t* var_sizes,
                            const int32_t* var_levels, int n_exact, const int64_t* const* ids_a,
                            const int64_t* const* ids_b, int64_t id_div, int64_t* ids, int64_t capacity,
                            int64_t* pattern_offsets, int64_t* n_union, fh_stream_t stream);

/*
 * fh_pattern_merge with a lookup aid for listed exact variables.
 *
 * The same merge and outputs as fh_pattern_merge (faster/comparison.py:
 * 695-725, faster/deduplication.py:773-803) over the same n_vars lists. For
 * a variable v with ids_a[v] and ids_b[v] non-null (device int64 per-row
 * value ids; a negative id is a row without a value), the caller states
 * that v has 2 levels and that its list is exactly the pairs with
 * ids_a[v][id / id_div] == ids_b[v][id % id_div] >= 0, at level 1 (the
 * exact-equality lists of faster/comparison.py:500-600 and
 * faster/deduplication.py:628-701; for deduplication ids_b[v] = ids_a[v]
 * and id_div = n). The other lists' entries then take v's level from the two
 * ids instead of a search of v's list; null entries are searched as usual.
 * Every pair id of every list must index both arrays (id / id_div below the
 * length of ids_a[v], id % id_div below that of ids_b[v]), as the lists of
 * those rows' pairs do.
 */
int fh_pattern_merge_listed(int n_vars, const uint64_t* const* var_keys, const int64_t* var_sizes,
                            const int32_t* var_levels, const int64_t* const* ids_a, const int64_t* const* ids_b,
                            int64_t id_div, int64_t* ids, int64_t capacity, int64_t* pattern_offsets,
                            int64_t* n_union, fh_stream_t stream);

/*
 * Fellegi-Sunter EM over a pattern histogram, FP64, one workgroup running
 * every iteration on the device.
 *
 * Replaces Estimation.fit / _match_probability (faster/estimation.py:72-158).
 * levels (host int32[n_vars]); counts (device int64[prod(levels)]) in
 * itertools.product order. params (device double): in = initial values, out =
 * final values, laid out [lambda, pi_0 of var 0..n-1 (levels each), pi_1 of
 * var 0..n-1]. ksi (device double[n_patterns]) receives the match probability
 * of every pattern under the final parameters. Stops when max |delta pi| <
 * tol or after max_iter iterations; *iterations and *converged (host) report
 * which.
 */
int fh_em_fit(int n_vars, const int32_t* levels, const int64_t* counts, double* params,
              double tol, int max_iter, double* ksi, int* iterations, int* converged,
              fh_stream_t stream);

/*
 * fh_em_fit without any allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops that re-fit EM on updated histograms, harnesses that
 * forbid hipMalloc). Arguments, results and kernels are those of fh_em_fit;
 * the per-pattern weights and the status words live in a library-owned
 * device array of 1 MiB per GPU (held until the process ends), so the
 * histogram may have at most 65,536 patterns (FH_ERR_ARG beyond). The call
 * allocates nothing. With iterations and converged both null it also reads
 * nothing back (no synchronisation); otherwise it synchronises the stream.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several streams or host threads) race and must be serialised by the
 * caller.
 */
int fh_em_fit_noalloc(int n_vars, const int32_t* levels, const int64_t* counts, double* params,
                      double tol, int max_iter, double* ksi, int* iterations, int* converged,
                      fh_stream_t stream);

/* Match probability of every pattern under `params` (layout as above);
 * faster/estimation.py:72-96 in FP64. */
int fh_match_probability(int n_vars, const int32_t* levels, const double* params, double* ksi,
                         fh_stream_t stream);

/*
 * Differentiable Fellegi-Sunter log-likelihood (FP64). No counterpart in the
 * reference, which only runs EM (faster/estimation.py:98-158); these give the
 * likelihood, its gradient and its Hessian over logits so that standard
 * errors and gradient-based fits are available.
 *
 * theta (device double[1 + 2 * sum(levels)]): [a, beta, alpha] with
 * lambda = sigmoid(a), pi_0[k] = softmax(beta_k), pi_1[k] = softmax(alpha_k)
 * (the layout [lambda, pi_0, pi_1] of fh_em_fit's params). counts: device int64[prod(levels)].
 *   fh_fs_loglik_fwd : *ll (device) = sum_p N_p log M_p; post (device,
 *                      optional) = posterior match probability per pattern.
 *   fh_fs_loglik_bwd : grad_theta = grad_out[0] * dLL/dtheta (hand-written).
 *   fh_fs_loglik_hess: hess (device double[dim * dim], dim = len(theta)) =
 *                      d2LL/dtheta2 in closed form.
 */
int fh_fs_loglik_fwd(int n_vars, const int32_t* levels, const int64_t* counts, const double* theta,
                     double* ll, double* post, fh_stream_t stream);
int fh_fs_loglik_bwd(int n_vars, const int32_t* levels, const int64_t* counts, const double* theta,
                     const double* grad_out, double* grad_theta, fh_stream_t stream);
int fh_fs_loglik_hess(int n_vars, const int32_t* levels, const int64_t* counts, const double* theta,
                      double* hess, fh_stream_t stream);

/*
 * Per-pair posterior gather: post[i] = ksi[k] for every pair i of pattern k,
 * following the pattern-major layout of fh_pattern_merge
 * (pattern_offsets device int64[n_patterns + 1]). The per-pair form of the
 * selection at faster/linkage.py:48-54.
 */
int fh_pair_posteriors(const int64_t* pattern_offsets, int n_patterns, const double* ksi,
                       double* post, int64_t n_pairs, fh_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* FASTER_HIP_H */
