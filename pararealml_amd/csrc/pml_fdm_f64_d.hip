// Fused stepping kernel instantiations: PrecF64, Cahn-Hilliard /
// convection-diffusion.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_D(fdm_table_f64_d, PrecF64)
}  // namespace pml
