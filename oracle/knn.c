/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the PLD post-processing of apps/pld/cuASR/knncuda.cu:170-237:
 * modified_insertion_sort_linear (per query column, insertion of each new
 * distance into the sorted first k rows, strict '>' shift so equal distances
 * keep index order) followed by compute_sqrt_linear (sqrt of the first k rows).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* dist: Q x R in the reference layout dist[q + r*Q]; it is sorted in place
 * exactly as the reference sorts it (only the first k rows are meaningful
 * afterwards); index receives [j*Q + q]. */
void oracle_knn_insertion_sort(float *dist, int32_t *index, int width, int height, int k) {
  for (int x = 0; x < width; ++x) {
    float *p_dist = dist + x;
    int32_t *p_index = index + x;
    p_index[0] = 0;
    for (int i = 1; i < height; ++i) {
      const float curr_dist = p_dist[(int64_t)i * width];
      const int curr_index = i;
      if (i >= k && curr_dist >= p_dist[(int64_t)(k - 1) * width]) continue;
      int j = i < k - 1 ? i : k - 1;
      while (j > 0 && p_dist[(int64_t)(j - 1) * width] > curr_dist) {
        p_dist[(int64_t)j * width] = p_dist[(int64_t)(j - 1) * width];
        p_index[(int64_t)j * width] = p_index[(int64_t)(j - 1) * width];
        --j;
      }
      p_dist[(int64_t)j * width] = curr_dist;
      p_index[(int64_t)j * width] = curr_index;
    }
  }
}

void oracle_knn_sqrt(float *dist, int width, int k) {
  for (int64_t i = 0; i < (int64_t)width * k; ++i) dist[i] = sqrtf(dist[i]);
}
