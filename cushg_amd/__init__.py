"""MI355X-native split-step Fourier SHG propagation (cuSHG's hot path)."""
