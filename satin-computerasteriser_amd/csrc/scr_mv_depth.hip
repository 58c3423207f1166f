// scr_mv_depth.hip — every shape of the batched depth-pass kernel (k_mv_depth), compiled
// apart from the other phases so that the library builds in parallel.
#include "scr_mv.hpp"

namespace scr {
namespace mv {

const KernelTable& depth_kernels() {
    static const KernelTable table = SCR_MV_TABLE(k_mv_depth);
    return table;
}

}  // namespace mv
}  // namespace scr
