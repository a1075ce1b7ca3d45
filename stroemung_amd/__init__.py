"""stroemung_amd: the stroemung NaSt2D time step on MI355X (gfx950).

The product is the C-ABI library libstroemung_hip.so (HIP kernels plus the
C++ host that owns device memory); this package is the Python host above it,
mirroring the reference's `Simulation` / `SimulationGrid` API.
"""
