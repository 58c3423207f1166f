// scr_mv_nearest.hip — every shape of the batched nearest-point kernel (k_mv_nearest), compiled
// apart from the other phases so that the library builds in parallel.
#include "scr_mv.hpp"

namespace scr {
namespace mv {

const KernelTable& nearest_kernels() {
    static const KernelTable table = SCR_MV_TABLE(k_mv_nearest);
    return table;
}

}  // namespace mv
}  // namespace scr
