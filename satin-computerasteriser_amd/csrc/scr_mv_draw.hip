// scr_mv_draw.hip — the batched draw kernel of one phase (SCR_MV_PHASE: 0
// depth, 1 colour, 2 nearest) in one shape (SCR_MV_SHAPE, an index into
// kShapes), for both footprint classes and both tile kinds.  The build
// compiles this file once per (phase, shape) pair.
#include "scr_mv.hpp"

#if !defined(SCR_MV_PHASE) || !defined(SCR_MV_SHAPE)
#error "compile with -DSCR_MV_PHASE=<0..2> -DSCR_MV_SHAPE=<0..5>"
#endif

namespace scr {
namespace mv {

#if SCR_MV_PHASE == 0
#define SCR_MV_KERNEL k_mv_depth
#elif SCR_MV_PHASE == 1
#define SCR_MV_KERNEL k_mv_color
#else
#define SCR_MV_KERNEL k_mv_nearest
#endif

#define SCR_MV_DEFINE(P, S)                                                                                 \
    const ShapeKernels& SCR_MV_ACCESSOR(P, S)() {                                                          \
        constexpr int kNT = kShapes[S].threads, kPPT = kShapes[S].ppt;                                     \
        static const ShapeKernels t = {{SCR_MV_KERNEL<false, false, kNT, kPPT>,                            \
                                        SCR_MV_KERNEL<false, true, kNT, kPPT>},                            \
                                       {SCR_MV_KERNEL<true, false, kNT, kPPT>,                             \
                                        SCR_MV_KERNEL<true, true, kNT, kPPT>}};                            \
        return t;                                                                                          \
    }
#define SCR_MV_DEFINE_EXPANDED(P, S) SCR_MV_DEFINE(P, S)
SCR_MV_DEFINE_EXPANDED(SCR_MV_PHASE, SCR_MV_SHAPE)

}  // namespace mv
}  // namespace scr
