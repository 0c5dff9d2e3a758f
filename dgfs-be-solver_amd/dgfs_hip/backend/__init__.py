from dgfs_hip.backend.base import HIPBackend, NotSuitableError  # noqa: F401
