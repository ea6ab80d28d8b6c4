// Fused stepping kernel instantiations: PrecF32, shallow water / wave.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_B(fdm_table_f32_b, PrecF32)
}  // namespace pml
