"""fvens_amd -- MI355X-native residual, Jacobian and preconditioner engine (package stub)."""
