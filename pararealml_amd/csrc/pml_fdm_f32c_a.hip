// Fused stepping kernel instantiations: PrecF32C, diffusion.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_A(fdm_table_f32c_a, PrecF32C)
}  // namespace pml
