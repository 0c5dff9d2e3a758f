"""dgfs_hip: MI355X (gfx950) right-hand-side hot path of the DGFS Boltzmann solver.

Host code keeps the frfs interfaces of the reference (scattering-model plug-ins,
backend kernel providers, DGFS system and steppers); the arithmetic runs in
hand-written HIP kernels behind the C-ABI of ``libdgfs_hip.so``.
"""
__version__ = "0.1.0"
