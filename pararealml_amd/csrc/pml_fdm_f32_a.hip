// Fused stepping kernel instantiations: PrecF32, diffusion.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_A(fdm_table_f32_a, PrecF32)
}  // namespace pml
