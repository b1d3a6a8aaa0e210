// Adjoint of one RSF step (placeholder until the kernels land).
#include "rsf_internal.hpp"

namespace rsf {

int adjoint_step(rsf_context*, const TilePlan&, const void*, int, const float*, const float*, float*, float*,
                 const rsf_params*, const Derived&, hipStream_t) {
    return set_error(RSF_ERR_UNSUPPORTED, "adjoint not built");
}

}  // namespace rsf
