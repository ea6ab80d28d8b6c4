// Fused stepping kernel instantiations: PrecF32C, Cahn-Hilliard /
// convection-diffusion.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_D(fdm_table_f32c_d, PrecF32C)
}  // namespace pml
