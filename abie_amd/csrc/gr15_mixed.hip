// gr15_mixed.hip -- GR15 engine instantiated for ABIE_FORCE_MIXED forces (one
// translation unit per force mode so the three compile in parallel).
#include "gr15_impl.hpp"

namespace abie {
hipError_t launch_gr15_mixed(int cfg, int ctrl, const GR15Args &a, int nb, size_t lds, hipStream_t st) {
    return launch_gr15_fm<ABIE_FORCE_MIXED>(cfg, ctrl, a, nb, lds, st);
}
}  // namespace abie
