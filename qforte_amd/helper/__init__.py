"""String-based construction helpers."""
