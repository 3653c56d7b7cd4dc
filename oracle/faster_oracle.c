/*
 * Oracle (test infrastructure only): a plain-C restatement of the reference's
 * per-pair device logic. Nothing in the product links or calls this file; only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, and
 * only as the checker.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * reference repository root). The loops keep the reference's structure
 * (per-pair bool work arrays, greedy first-unused match, the half-count of
 * transpositions, the fp32 formula with one fp64 divide by 3.0).
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off so that the only fused
 * multiply-add is the explicit fmaf below, which mirrors the contraction the
 * reference's NVRTC build applies to `dist += p * prefix * (1 - dist)`).
 */
#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Restates faster/comparison.py:11-131 (identical copy at
 * faster/deduplication.py:10-130). `use_fma` selects between the contracted
 * final update (fmaf, what NVRTC emits under its default --fmad=true) and the
 * uncontracted form; both are exposed so tests can show the level is
 * insensitive to the choice. */
static float jaro_winkler(const unsigned char *str1, int len1, bool *hash_str1,
                          const unsigned char *str2, int len2, bool *hash_str2,
                          float p, int use_fma) {
  if (len1 == 0 || len2 == 0) {
    return 0.0f; /* comparison.py:34-37 */
  }
  int max_dist = (len1 > len2 ? len1 : len2) / 2 - 1; /* :44, may be -1 */
  float match = 0;
  for (int i = 0; i < len1; i++) { /* :48-64 */
    int jlo = i - max_dist > 0 ? i - max_dist : 0;
    int jhi = len2 < i + max_dist + 1 ? len2 : i + max_dist + 1;
    for (int j = jlo; j < jhi; j++) {
      if (str1[i] == str2[j] && hash_str2[j] == false) {
        hash_str1[i] = true;
        hash_str2[j] = true;
        match++;
        break;
      }
    }
  }
  if (match == 0) {
    return 0.0f; /* :66-69 */
  }
  float t = 0; /* :73-99 */
  int point = 0;
  for (int i = 0; i < len1; i++) {
    if (hash_str1[i] == true) {
      while (hash_str2[point] == false) {
        point++;
      }
      if (str1[i] != str2[point++]) {
        t++;
      }
    }
  }
  t /= 2;
  /* :102 -- three fp32 divides, fp32 sum (left to right), fp64 divide by 3.0,
   * rounded back to fp32 on assignment. */
  float dist = (float)((double)((match / (float)len1) + (match / (float)len2) +
                                ((match - t) / match)) /
                       3.0);
  float prefix = 0; /* :106-120 */
  int lim = len1 < len2 ? len1 : len2;
  if (lim > 4) lim = 4;
  for (int i = 0; i < lim; i++) {
    if (str1[i] == str2[i]) {
      prefix++;
    } else {
      break;
    }
  }
  /* :123 */
  if (use_fma) {
    dist = fmaf(p * prefix, 1 - dist, dist);
  } else {
    float q = (p * prefix) * (1 - dist);
    dist = dist + q;
  }
  return dist;
}

/* Score of one pair, allocating its own zeroed work arrays (the reference's
 * cp.zeros buffers, comparison.py:305-307). */
float oracle_jw_score(const unsigned char *s1, int len1, const unsigned char *s2,
                      int len2, float p, int use_fma) {
  bool h1[4096];
  bool h2[4096];
  if (len1 > 4096 || len2 > 4096) return -1.0f;
  memset(h1, 0, (size_t)len1);
  memset(h2, 0, (size_t)len2);
  return jaro_winkler(s1, len1, h1, s2, len2, h2, p, use_fma);
}

/* The final formula alone, for a given (len1, len2, m, tt, prefix): the same
 * arithmetic as lines :99-123 once m, the mismatch count tt (t = tt/2) and the
 * prefix are known. Used to enumerate every tuple. */
float oracle_jw_formula(int len1, int len2, int m, int tt, int prefix_i, float p,
                        int use_fma) {
  if (len1 == 0 || len2 == 0 || m == 0) return 0.0f;
  float match = (float)m;
  float t = (float)tt;
  t /= 2;
  float dist = (float)((double)((match / (float)len1) + (match / (float)len2) +
                                ((match - t) / match)) /
                       3.0);
  float prefix = (float)prefix_i;
  if (use_fma) {
    dist = fmaf(p * prefix, 1 - dist, dist);
  } else {
    float q = (p * prefix) * (1 - dist);
    dist = dist + q;
  }
  return dist;
}

/* Level of a score under the reference's thresholding (comparison.py:321-329):
 * level 1 when lo <= s < hi, level 2 when s >= hi, compared in fp32. */
static inline int level_of(float s, float lo, float hi) {
  if (s >= hi) return 2;
  if (s >= lo && s < hi) return 1;
  return 0;
}

/* All-pairs linkage scores for rows [r0, r1) of set 1 against all of set 2,
 * thresholded into two levels; ids are idx * n2 + idy in the order the
 * reference's dense output and argwhere produce them (ascending). Restates
 * the kernel faster/comparison.py:133-186 and the thresholding :321-329.
 * Returns the number of (level-1, level-2) hits through n_out; hits beyond
 * the capacities are counted but not written. */
void oracle_jw_levels_link(const unsigned char *str1, const int64_t *off1,
                           int64_t r0, int64_t r1, const unsigned char *str2,
                           const int64_t *off2, int64_t n2, float p, float lo,
                           float hi, int use_fma, int64_t *out1, int64_t cap1,
                           int64_t *out2, int64_t cap2, int64_t *n_out) {
  int64_t c1 = 0, c2 = 0;
  int maxl2 = 0;
  for (int64_t j = 0; j < n2; j++) {
    int l = (int)(off2[j + 1] - off2[j]);
    if (l > maxl2) maxl2 = l;
  }
  int maxl1 = 0;
  for (int64_t i = r0; i < r1; i++) {
    int l = (int)(off1[i + 1] - off1[i]);
    if (l > maxl1) maxl1 = l;
  }
  bool *h1 = (bool *)calloc((size_t)maxl1 + 1, 1);
  bool *h2 = (bool *)calloc((size_t)maxl2 + 1, 1);
  for (int64_t idx = r0; idx < r1; idx++) {
    const unsigned char *s1 = str1 + off1[idx];
    int len1 = (int)(off1[idx + 1] - off1[idx]);
    for (int64_t idy = 0; idy < n2; idy++) {
      const unsigned char *s2 = str2 + off2[idy];
      int len2 = (int)(off2[idy + 1] - off2[idy]);
      memset(h1, 0, (size_t)len1);
      memset(h2, 0, (size_t)len2);
      float s = jaro_winkler(s1, len1, h1, s2, len2, h2, p, use_fma);
      int lvl = level_of(s, lo, hi);
      int64_t id = idx * n2 + idy;
      if (lvl == 1) {
        if (c1 < cap1) out1[c1] = id;
        c1++;
      } else if (lvl == 2) {
        if (c2 < cap2) out2[c2] = id;
        c2++;
      }
    }
  }
  free(h1);
  free(h2);
  n_out[0] = c1;
  n_out[1] = c2;
}

/* Deduplication scores for unique rows [r0, r1) against columns col <= row of
 * the same set (the lower triangle), restating the kernel
 * faster/deduplication.py:132-196: string1 is the row value, string2 the
 * column value, and the diagonal row == col scores exactly 1. Ids are
 * row * n + col, ascending. */
void oracle_jw_levels_dedup(const unsigned char *str, const int64_t *off,
                            int64_t n, int64_t r0, int64_t r1, float p, float lo,
                            float hi, int use_fma, int64_t *out1, int64_t cap1,
                            int64_t *out2, int64_t cap2, int64_t *n_out) {
  int64_t c1 = 0, c2 = 0;
  int maxl = 0;
  for (int64_t j = 0; j < n; j++) {
    int l = (int)(off[j + 1] - off[j]);
    if (l > maxl) maxl = l;
  }
  bool *h1 = (bool *)calloc((size_t)maxl + 1, 1);
  bool *h2 = (bool *)calloc((size_t)maxl + 1, 1);
  for (int64_t row = r0; row < r1; row++) {
    const unsigned char *s1 = str + off[row];
    int len1 = (int)(off[row + 1] - off[row]);
    for (int64_t col = 0; col <= row; col++) {
      float s;
      if (row != col) {
        const unsigned char *s2 = str + off[col];
        int len2 = (int)(off[col + 1] - off[col]);
        memset(h1, 0, (size_t)len1);
        memset(h2, 0, (size_t)len2);
        s = jaro_winkler(s1, len1, h1, s2, len2, h2, p, use_fma);
      } else {
        s = 1.0f; /* deduplication.py:185-190 */
      }
      int lvl = level_of(s, lo, hi);
      int64_t id = row * n + col;
      if (lvl == 1) {
        if (c1 < cap1) out1[c1] = id;
        c1++;
      } else if (lvl == 2) {
        if (c2 < cap2) out2[c2] = id;
        c2++;
      }
    }
  }
  free(h1);
  free(h2);
  n_out[0] = c1;
  n_out[1] = c2;
}

/* Membership mask of sorted arr1 elements in sorted arr2, restating the
 * binary search kernel faster/search.py:8-50 (in = 1: intersect mask,
 * in = 0: set-difference mask). */
void oracle_binary_search_mask(const int64_t *arr1, int64_t n1, const int64_t *arr2,
                               int64_t n2, int in, uint8_t *out) {
  for (int64_t id = 0; id < n1; id++) {
    int64_t x = arr1[id];
    int64_t low = 0, high = n2 - 1;
    bool found = false;
    while (low <= high) {
      int64_t mid = (low + high) / 2;
      if (arr2[mid] == x) {
        found = true;
        break;
      }
      if (arr2[mid] < x) {
        low = mid + 1;
      } else {
        high = mid - 1;
      }
    }
    out[id] = (uint8_t)(in ? found : !found);
  }
}
