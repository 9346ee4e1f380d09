// gr15_fp64_massive.hip -- GR15 engine for ABIE_FORCE_FP64 forces under the ABIE_CTRL_MASSIVE controller
// (one translation unit per pair so that they compile in parallel).
#include "gr15_quad.hpp"

namespace abie {
hipError_t launch_gr15_fp64_massive(int cfg, const GR15Args &a, int nb, size_t lds, hipStream_t st, bool coop) {
    return launch_gr15_fm_ct<ABIE_FORCE_FP64, ABIE_CTRL_MASSIVE>(cfg, a, nb, lds, st, coop);
}
hipError_t launch_gr15q_fp64_massive(int q, const GR15Args &a, int nb, size_t lds, hipStream_t st) {
    return launch_gr15q_fm_ct<ABIE_FORCE_FP64, ABIE_CTRL_MASSIVE>(q, a, nb, lds, st);
}
}  // namespace abie
