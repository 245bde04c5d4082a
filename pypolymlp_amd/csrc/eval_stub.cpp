// Evaluation and ridge entry points not yet available on the device.
#include "capi_internal.hpp"

namespace plm {

void upload_potential(ModelHandle&) {}

void eval_single(ModelHandle&, const double*, const double*, const int*, int, bool, double*,
                 double*, double*) {
    throw std::runtime_error("plm_eval: not available in this build");
}

void eval_batch(ModelHandle&, BatchHandle&, bool, double*, double*, double*) {
    throw std::runtime_error("plm_eval_batch: not available in this build");
}

void ridge_solve(int, const double*, const double*, const double*, int, double*, int*) {
    throw std::runtime_error("plm_ridge: not available in this build");
}

}  // namespace plm
