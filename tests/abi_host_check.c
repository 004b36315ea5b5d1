/* Host-side paths of the C-ABI, run without a GPU under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_abi_sanitizers.py builds the library's
 * host code with both and this driver against it): the sizing functions, the
 * host tile packer over synthetic records with short, empty and over-long rows,
 * and the argument checks of the device entry points, which must reject null
 * or inconsistent arguments with an error status and a message before any
 * device work.  Prints "ALL OK" and exits 0 when every check holds. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dips_hip.h"

static int failures = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);    \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static int64_t net_params(int L, int H) { return 4 * H + (int64_t)(L - 1) * (H * H + H) + H + 1; }

static void sizing(void) {
  dips_net_t net;
  memset(&net, 0, sizeof net);
  net.Lm = 3, net.Hm = 64, net.Lp = 3, net.Hp = 64;
  CHECK(dips_version() == 1);
  CHECK(dips_param_count(&net) == 2 * net_params(3, 64));
  CHECK(dips_mfma_supported(&net) == 1);
  const int64_t w = dips_mfma_workspace(&net, 262144, 2000000);
  CHECK(w > 2000000 + 262144);
  dips_net_t small = net;
  small.Lm = 1, small.Hm = 1, small.Lp = 2, small.Hp = 10;
  CHECK(dips_param_count(&small) == net_params(1, 1) + net_params(2, 10));
  CHECK(dips_partial_size(&small, 65536) > 0);
  CHECK(dips_tile_partial_size(&small) > 0);
  /* invalid networks and sizes */
  dips_net_t bad = net;
  bad.Lm = 0;
  CHECK(dips_param_count(&bad) == -1);
  CHECK(dips_param_count(NULL) == -1);
  CHECK(dips_partial_size(&small, 0) == -1);
  CHECK(dips_mfma_workspace(&net, 0, 10) == -1);
  CHECK(dips_net_supported(&bad) == 0);
  bad = net;
  bad.act = 7;
  CHECK(dips_param_count(&bad) == -1);
  CHECK(strlen(dips_last_error()) > 0);
  /* hash network sizes */
  dips_hashgrid_t hg;
  memset(&hg, 0, sizeof hg);
  hg.levels = 16, hg.features = 2, hg.table_size = 1 << 19, hg.n_min = 16, hg.n_max = 2048, hg.bound = 1.0f;
  const int64_t rows = dips_hash_table_rows(&hg);
  CHECK(rows > 0 && rows <= 16 * (int64_t)(1 << 19));
  CHECK(dips_hash_param_count(&hg) > rows * 2);
  int64_t doubles = 0;
  CHECK(dips_hash_workspace(&hg, 65536, 500000, &doubles) > 0 && doubles > 0);
  hg.levels = 1; /* outside 2..16 */
  CHECK(dips_hash_table_rows(&hg) < 0);
}

/* synthetic CSR row offsets: lengths 0..90 (some rows longer than a tile) */
static void tiles(void) {
  enum { NR = 5000, NB = 3 };
  int32_t* row_ptr = malloc(sizeof(int32_t) * (NR + 1));
  uint32_t s = 12345u;
  row_ptr[0] = 0;
  for (int r = 0; r < NR; ++r) {
    s = s * 1664525u + 1013904223u;
    int len = (int)((s >> 8) % 100u);
    len = len < 70 ? len % 9 : len - 10; /* mostly short rows, some empty, some 60..89 */
    row_ptr[r + 1] = row_ptr[r] + len;
  }
  const int64_t begin[NB] = {0, 1700, 3300}, nrows[NB] = {1700, 1600, 1700};
  int64_t tb[NB + 1];
  const int64_t k = dips_pack_tiles(row_ptr, NR, begin, nrows, NB, NULL, 0, tb);
  CHECK(k > 0 && tb[NB] == k);
  int32_t* tile_row = malloc(sizeof(int32_t) * (size_t)k);
  int64_t tb2[NB + 1];
  CHECK(dips_pack_tiles(row_ptr, NR, begin, nrows, NB, tile_row, k, tb2) == k);
  for (int b = 0; b <= NB; ++b) CHECK(tb[b] == tb2[b]);
  /* each batch's boundaries: start at its first row, end at its end, increase, and a tile of several rows
     holds at most 64 entries and 64 rows */
  for (int b = 0; b < NB; ++b) {
    CHECK(tile_row[tb[b]] == begin[b]);
    CHECK(tile_row[tb[b + 1] - 1] == begin[b] + nrows[b]);
    for (int64_t i = tb[b]; i + 1 < tb[b + 1]; ++i) {
      const int32_t r0 = tile_row[i], r1 = tile_row[i + 1];
      CHECK(r1 > r0);
      if (r1 - r0 > 1) {
        CHECK(r1 - r0 <= 64);
        int64_t slots = 0;
        for (int32_t r = r0; r < r1; ++r) slots += row_ptr[r + 1] - row_ptr[r] > 0 ? row_ptr[r + 1] - row_ptr[r] : 1;
        CHECK(slots <= 64);
      }
    }
  }
  /* capacity too small, decreasing offsets and batches outside the records are errors */
  CHECK(dips_pack_tiles(row_ptr, NR, begin, nrows, NB, tile_row, k - 1, tb2) == -1);
  const int64_t outside[1] = {4990}, outside_n[1] = {20};
  CHECK(dips_pack_tiles(row_ptr, NR, outside, outside_n, 1, NULL, 0, NULL) == -1);
  row_ptr[10] = row_ptr[9] - 1;
  CHECK(dips_pack_tiles(row_ptr, NR, begin, nrows, NB, NULL, 0, NULL) == -1);
  CHECK(dips_pack_tiles(NULL, NR, begin, nrows, NB, NULL, 0, NULL) == -1);
  free(tile_row);
  free(row_ptr);
}

/* device entry points reject null or inconsistent arguments before any device work */
static void argument_checks(void) {
  dips_net_t net;
  memset(&net, 0, sizeof net);
  net.Lm = 3, net.Hm = 64, net.Lp = 3, net.Hp = 64;
  dips_opt_t opt;
  memset(&opt, 0, sizeof opt);
  dips_grid_t grid;
  memset(&grid, 0, sizeof grid);
  dips_records_t rec;
  memset(&rec, 0, sizeof rec);
  float f = 0.0f;
  int32_t i32 = 0;
  double d = 0.0;
  int32_t it = 0;
  CHECK(dips_mfma_train_step(NULL, NULL, NULL, &net, &opt, 1.0f, &f, &f, &f, &i32, &f, 1, &f, 1, &d, 1, NULL) != 0);
  CHECK(dips_mfma_train_step(&grid, &rec, NULL, &net, &opt, 1.0f, &f, &f, &f, &i32, &f, 1, &f, 1, &d, 1, NULL) != 0);
  CHECK(strlen(dips_last_error()) > 0);
  CHECK(dips_mfma_loss_grad(NULL, NULL, NULL, &net, 1.0f, &f, &f, &d, &f, 1, &d, 1, NULL) != 0);
  CHECK(dips_loss_grad(NULL, NULL, 0, 1, 0, &net, &f, &f, 1, 1, &i32, NULL) != 0);
  CHECK(dips_tile_train_step(NULL, NULL, NULL, 0, 1, &net, &opt, 1.0f, &f, &f, &f, &i32, &f, 1, &f, 1, NULL, NULL) !=
        0);
  CHECK(dips_bicgstab(NULL, NULL, &f, &f, &d, 10, 1e-6, &it, &d, NULL) != 0);
  CHECK(dips_grid_apply(NULL, NULL, &f, &f, 1, NULL) != 0);
  CHECK(dips_eval(NULL, &f, &f, NULL, NULL, 1, &f, NULL, NULL, NULL) != 0);
  CHECK(dips_geom_assemble(NULL, NULL) != 0);
  CHECK(dips_hash_train_step(NULL, NULL, NULL, NULL, &opt, 1.0f, &f, &f, &f, &i32, &f, 1, &f, 1, &d, 1, NULL) !=
        0);
  CHECK(dips_batch_workspace(NULL, 0, &net) < 0);
  /* a points-only set (no axes) needs extra points; axes must be all set or all empty */
  dips_grid_t pts;
  memset(&pts, 0, sizeof pts);
  dips_records_t r1;
  memset(&r1, 0, sizeof r1);
  r1.row_point = &i32, r1.row_ptr = &i32, r1.row_c = &f, r1.nrows = 1;
  CHECK(dips_loss_grad(&pts, &r1, 0, 1, 0, &net, &f, &f, 1, 1, &i32, NULL) != 0);
  CHECK(strstr(dips_last_error(), "invalid grid") != NULL);
  pts.ny = 1;
  pts.extra = &f, pts.nextra = 1;
  CHECK(dips_loss_grad(&pts, &r1, 0, 1, 0, &net, &f, &f, 1, 1, &i32, NULL) != 0);
  CHECK(strstr(dips_last_error(), "invalid grid") != NULL);
}

int main(void) {
  sizing();
  tiles();
  argument_checks();
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("ALL OK\n");
  return 0;
}
