// Fused stepping kernel instantiations: PrecF32, Cahn-Hilliard /
// convection-diffusion.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_D(fdm_table_f32_d, PrecF32)
}  // namespace pml
