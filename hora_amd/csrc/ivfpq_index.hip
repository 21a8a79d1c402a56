// ivfpq_index.hip -- IVFPQIndex on the device (pq_idx.rs:267-523).
#include "index.h"

namespace hm {

std::unique_ptr<Index> make_ivfpq(uint64_t dim, const hm_params &p)
{
    (void)dim;
    (void)p;
    fail(HM_ERR_NOT_IMPLEMENTED, "IVFPQIndex is not available in this build yet");
}

}  // namespace hm
