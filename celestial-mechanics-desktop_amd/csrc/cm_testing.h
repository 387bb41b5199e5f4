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
 * build (engine/src/cuda/gpu_tree.cu:269-464); and the target groups with
 * the order in which the traversal takes them.
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

/* Builds the tree and the target groups for the current device state, as a
 * force evaluation does before its traversal, and copies out:
 *   groups[2 * g]  sorted body range [start, end) of each group, in body order;
 *   queue[4 * k]   the traversal's queue entries (group start, group end,
 *                  targets start, targets end): queue q holds entries
 *                  [qbounds[q], qbounds[8 + q]), the parts of its fat groups
 *                  (bit-exact FP64 only), then its heavy groups (at least
 *                  twice the mean work of the last traversal), then the rest;
 *   qbounds[16]    the queues' first entries, then their ends;
 *   cost[count]    per body (caller's index): the work units of its group in
 *                  the last traversal, which set that order (0 before any);
 *   order[count]   original index of each sorted position;
 *   *n_groups      the number of groups.
 * At most cap groups and cap queue entries are copied; any output pointer may
 * be NULL. */
int32_t cm_test_groups(void* engine, int32_t cap, int32_t* groups, int32_t* queue,
                       int32_t* qbounds, int32_t* cost, int32_t* order, int32_t* n_groups);

#ifdef __cplusplus
}
#endif
#endif /* CM_TESTING_H */
