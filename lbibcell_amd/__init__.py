"""MI355X-native per-timestep lattice path of an LBIBCell-compatible tissue
simulator: D2Q9 + D2Q5 lattice Boltzmann with immersed-boundary membranes.

The compute path is liblbib_hip.so (hand-written HIP for gfx950 behind the C
ABI in include/lbib.h); the host runtime that mirrors the reference's
SimulationRunner / ForceSolver / BioSolver API is liblbib_host.so.
"""
__all__ = ["capi"]
