// mt_fused_q27_f16.hip -- the fused update kernels of D3Q27 with _Float16 storage.
#include "mt_fused.h"

namespace mt {
MT_FUSED_INSTANTIATE(27, _Float16)
}  // namespace mt
