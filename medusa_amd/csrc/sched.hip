// Feature encoding of trace records for the scheduler MLP (C ABI entry point; the
// encoders themselves are host/device inline functions in sched.h).
#include "sched.h"

namespace medusa {

void Sched::policy_row(uint32_t kind, const double* r, double action, float* x) {
    if (kind == MEDUSA_POLICY_DIRECTION)
        direction_row(r, action, x);
    else if (kind == MEDUSA_POLICY_DELTA)
        delta_row(r, action, x);
    else
        variant_row(r, action, x);
}

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_policy_features(uint32_t kind, const double* rec, double action, float* row) {
    return guard([&] {
        MDS_CHECK(rec && row, MEDUSA_ERR_ARG, "policy_features: null argument");
        MDS_CHECK(kind == MEDUSA_POLICY_DIRECTION || kind == MEDUSA_POLICY_DELTA || kind == MEDUSA_POLICY_VARIANT,
                  MEDUSA_ERR_ARG,
                  "policy_features: unknown kind %u", kind);
        Sched::policy_row(kind, rec, action, row);
    });
}
