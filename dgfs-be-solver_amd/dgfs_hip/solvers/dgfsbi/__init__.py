"""Binary (multi-species) DGFS solver pieces on the HIP path (frfs/solvers/dgfsbi)."""
