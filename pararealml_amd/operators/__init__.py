"""Solver operators: FDM (fused gfx950 kernels) and Parareal."""
