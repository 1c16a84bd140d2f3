/* A plain C caller of libvxl through include/vxl.h only (what an FFI binding sees): one
 * triangle voxelized densely and into a tree, the tree downloaded, rays traversed, and an
 * argument error reported by status and message. Prints "client ok" and exits 0 on success.
 * Usage: client (needs a GPU). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vxl.h"

#define CHECK(call)                                                                     \
  do {                                                                                  \
    vxl_status s_ = (call);                                                             \
    if (s_ != VXL_OK) {                                                                 \
      fprintf(stderr, "%s failed: %s (%s)\n", #call, vxl_status_str(s_), vxl_last_error()); \
      return 1;                                                                         \
    }                                                                                   \
  } while (0)

int main(void) {
  vxl_ctx *ctx = NULL;
  CHECK(vxl_ctx_create(0, &ctx));
  /* the reference's hand triangle (voxelizer/tests/differential.rs:193-221) at n = 8 */
  const float tri[9] = {0.5f, 0.5f, 0.5f, 4.5f, 1.0f, 1.2f, 1.1f, 4.2f, 2.0f};
  vxl_grid g;
  memset(&g, 0, sizeof(g));
  g.n = 8;
  g.voxel_size = 1.0f;
  uint32_t words[16];
  uint64_t pairs = 0;
  CHECK(vxl_voxelize_dense(ctx, tri, 1, &g, 1e-4f, words, &pairs));
  int set = 0;
  for (int i = 0; i < 16; i++) set += __builtin_popcount(words[i]);
  const int has111 = (words[(1 + 8 * (1 + 8 * 1)) >> 5] >> ((1 + 8 * (1 + 8 * 1)) & 31)) & 1;
  vxl_tree *tree = NULL;
  CHECK(vxl_voxelize_tree(ctx, tri, 1, &g, 1e-4f, &tree, NULL));
  vxl_tree_info info;
  CHECK(vxl_tree_info_get(tree, &info));
  uint64_t *leaves = calloc(info.leaf_count * 8 + 8, 8);
  CHECK(vxl_tree_download(tree, NULL, leaves, NULL, NULL));
  int tree_set = 0;
  for (uint64_t i = 0; i < info.leaf_count * 8; i++) tree_set += __builtin_popcountll(leaves[i]);
  /* a ray down the x axis through voxel row (y, z) = (1, 1) */
  const float rays[6] = {-3.0f, 1.5f, 1.5f, 1.0f, 0.0f, 0.0f};
  uint32_t hit[4];
  double t = 0;
  CHECK(vxl_traverse(ctx, tree, rays, 1, hit, &t));
  /* an argument error comes back as a status and a message */
  vxl_grid bad = g;
  bad.n = 64; /* not 8 * 4^k */
  const vxl_status es = vxl_voxelize_dense(ctx, tri, 1, &bad, 1e-4f, words, NULL);
  const int err_ok = es == VXL_NOT_REPRESENTABLE && strlen(vxl_last_error()) > 0;
  printf("pairs %llu set %d tree_set %d leaves %llu has111 %d hit %u %u %u %u t %.3f err %s\n",
         (unsigned long long)pairs, set, tree_set, (unsigned long long)info.leaf_count, has111,
         hit[0], hit[1], hit[2], hit[3], t, vxl_status_str(es));
  vxl_tree_free(tree);
  vxl_ctx_destroy(ctx);
  free(leaves);
  if (set == 0 || set != tree_set || !has111 || hit[3] != 1 || !err_ok) return 2;
  printf("client ok\n");
  return 0;
}
