/*
 * cm_testing.h -- test-only exports of libcmhip.so (not part of the public
 * C ABI in include/cmhip.h; no caller outside the parity tests needs them).
 *
 * They expose the intermediate results of the Barnes-Hut build so that the
 * tests can compare them bit for bit with the CPU restatement of the same
 * definitions: the Morton keys and stable sort order that replace the
 * reference engine's morton_code_kernel and radix sort
 * (engine/src/cuda/morton_kernel.cu:11-45,185-230), and the level-order
 * octree with its masses and centres of mass that replaces its insertion
 * build (engine/src/cuda/gpu_tree.cu:269-464).
 */
#ifndef CM_TESTING_H
#define CM_TESTING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Builds the tree for the current device state with the configured
 * morton_levels and leaf_capacity (bbox, keys, radix sort, gather, octree,
 * bottom-up mass) and copies out:
 *   keys[count]          sorted keys (inactive bodies last, with a sentinel);
 *   order[count]         original index of each sorted position;
 *   box4[4]              root cube centre x, y, z and half size;
 *   *n_active, *n_nodes  active bodies and nodes of the tree;
 *   info[5 * nodes]      per node: first child (-1 for a leaf), child count,
 *                        level, first sorted body, body count;
 *   com_mass[4 * nodes]  centre of mass x, y, z and mass;
 *   level_start[24]      first node of each level.
 * At most node_cap nodes are copied; any output pointer may be NULL.
 * Returns CM_ERR_POOL_EXHAUSTED when the tree outgrew the node pool. */
int32_t cm_test_tree(void* engine, int32_t node_cap, uint64_t* keys, int32_t* order,
                     double* box4, int32_t* n_active, int32_t* n_nodes, int32_t* info,
                     double* com_mass, int32_t* level_start);

#ifdef __cplusplus
}
#endif
#endif /* CM_TESTING_H */
