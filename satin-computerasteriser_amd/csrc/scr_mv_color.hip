// scr_mv_color.hip — every shape of the batched colour-pass kernel (k_mv_color), compiled
// apart from the other phases so that the library builds in parallel.
#include "scr_mv.hpp"

namespace scr {
namespace mv {

const KernelTable& color_kernels() {
    static const KernelTable table = SCR_MV_TABLE(k_mv_color);
    return table;
}

}  // namespace mv
}  // namespace scr
