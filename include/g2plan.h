/* include/g2plan.h -- query plan tables shared by libg2host.so (builder) and
 * libg2hip.so (matching kernels).
 *
 * The layout mirrors the arrays of the reference's Plan (system/plan.h:381-410)
 * and GMContext (system/appbase.h:11-68), bounded to 8 query vertices (the
 * reference's per-warp LDS partial match is [8] words, app_gmatch/gmatch.h:87).
 * Index conventions are the reference's:
 *   - by POSITION in the match order: bn*, shared, pre_bn*, after_bn*, strategy
 *   - by QUERY VERTEX id:             cond*, pre_cond*, after_cond*
 * Condition entries are (op, query vertex j): op 0 = LESS_THAN (M(u) < M(j)),
 * 1 = LARGER_THAN (M(u) > M(j)), 2 = NON_EQUAL (common/meta.h:18).
 * strategy: 0 = EXPAND, 1 = PREFIX, 2 = COUNT (common/meta.h:20).
 */
#ifndef G2PLAN_H
#define G2PLAN_H

#include <stdint.h>

#define G2_MAXQ 8

#ifdef __cplusplus
extern "C" {
#endif

typedef struct g2_plan {
    uint32_t n;                               /* query vertices, 2..8 */
    uint32_t root_degree;                     /* Plan::root_degree_ */
    int32_t  hop;                             /* Plan::GetHop() */
    uint32_t seq[G2_MAXQ];                    /* position -> query vertex (seq_) */
    uint32_t rev[G2_MAXQ];                    /* query vertex -> position (reverse_seq_) */
    uint32_t bn_count[G2_MAXQ];               /* backNeighborCountHost */
    uint32_t bn[G2_MAXQ][G2_MAXQ];            /* backNeighborsHost */
    uint32_t cond_count[G2_MAXQ];             /* condNumHost */
    uint32_t cond[G2_MAXQ][G2_MAXQ][2];       /* condOrderHost */
    uint32_t shared[G2_MAXQ];                 /* share_intersection */
    uint32_t pre_bn_count[G2_MAXQ];
    uint32_t pre_bn[G2_MAXQ][G2_MAXQ];
    uint32_t after_bn_count[G2_MAXQ];
    uint32_t after_bn[G2_MAXQ][G2_MAXQ];
    uint32_t pre_cond_count[G2_MAXQ];
    uint32_t pre_cond[G2_MAXQ][G2_MAXQ][2];
    uint32_t after_cond_count[G2_MAXQ];
    uint32_t after_cond[G2_MAXQ][G2_MAXQ][2];
    uint32_t strategy[G2_MAXQ + 1];
    uint32_t moving_lvl[G2_MAXQ];
} g2_plan;

#ifdef __cplusplus
}
#endif
#endif
