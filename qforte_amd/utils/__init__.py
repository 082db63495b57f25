"""Hot-path utilities: Pauli exponentials, Trotter products, reference states."""
