"""Flux-reconstruction kernels of the HIP path (bindings registered here)."""
from dgfs_hip import _lib  # noqa: F401
