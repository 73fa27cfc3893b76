"""ressim_amd: MI355X-native sparse Krylov path for the rs reservoir simulator.

The compute path is librsgpu.so (HIP, gfx950) behind the C ABI in
include/rsgpu.h; ``ressim_amd.mtj`` mirrors MTJ's solver/preconditioner
interfaces over it and ``ressim_amd.problems`` generates the seeded inputs.
"""
__all__ = ["mtj", "problems", "build"]
