// k_collide.hip -- collision detection, MergeOnly resolution and compaction.
#include <hip/hip_runtime.h>

#include "cm_internal.h"

namespace cm {

int collide_detect(Engine& e, cm_collision_event*, int32_t, int32_t* ne) {
    *ne = 0;
    return set_error(e, CM_ERR_INVALID_ARGUMENT, "collisions not available");
}
int collide_resolve(Engine& e, int32_t*) {
    return set_error(e, CM_ERR_INVALID_ARGUMENT, "collisions not available");
}
int collide_compact(Engine& e, int32_t*, int32_t*) {
    return set_error(e, CM_ERR_INVALID_ARGUMENT, "compaction not available");
}
void collide_free(Engine&) {}

}  // namespace cm
