// Reduced-precision (MFMA) reciprocal contraction of the Ewald table.
#include "ewald_common.hpp"

namespace pycd {

int launch_table_mfma(pycd_ctx* ctx, int precision, const double*, int,
                      const double*, const double*, int, const MicGeometry&,
                      const EwaldScalars&, double*, unsigned long long*,
                      hipEvent_t) {
  return fail(ctx, PYCD_ERR_ARG, "precision %d not built", precision);
}

}  // namespace pycd
