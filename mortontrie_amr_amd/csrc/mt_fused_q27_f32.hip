// mt_fused_q27_f32.hip -- the fused update kernels of D3Q27 with float storage.
#include "mt_fused.h"

namespace mt {
MT_FUSED_INSTANTIATE(27, float)
}  // namespace mt
