// LDS-staged flux_div for the hot flux combinations (placeholder: the
// dispatcher falls back to the generic kernel when this returns false).
#include "spd_internal.hpp"

namespace spd {

bool launch_flux_div_fast(const Grid&, const double*, double*, const spd_flux_desc&, bool, hipStream_t)
{
    return false;
}

} // namespace spd
