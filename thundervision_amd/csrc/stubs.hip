// Entry points whose kernels are not built yet: they fail loudly.
#include "kernels.hpp"

using namespace tdvh;

#define NOT_YET(name)                                                                              \
    do {                                                                                           \
        set_error("%s is not implemented in this build", name);                                   \
        return TDVH_ERR_UNSUPPORTED;                                                                \
    } while (0)

extern "C" int tdvh_batch_create(tdvh_ctx *, const tdvh_pair *, int, const tdvh_sgm_params *,
                                 tdvh_batch **) { NOT_YET("tdvh_batch_create"); }
extern "C" int tdvh_batch_destroy(tdvh_batch *) { NOT_YET("tdvh_batch_destroy"); }
extern "C" int tdvh_batch_config(tdvh_batch *, int, int, int) { NOT_YET("tdvh_batch_config"); }
extern "C" int tdvh_batch_run(tdvh_batch *, int, void *) { NOT_YET("tdvh_batch_run"); }
extern "C" int64_t tdvh_batch_cells(const tdvh_batch *) { return 0; }
extern "C" size_t tdvh_batch_scratch_bytes(const tdvh_batch *) { return 0; }
