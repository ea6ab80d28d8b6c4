// Fused stepping kernel instantiations: PrecF64, diffusion.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_A(fdm_table_f64_a, PrecF64)
}  // namespace pml
