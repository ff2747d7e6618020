"""MI355X-native per-timestep core of a Fourier-Fourier-Chebyshev channel DNS."""
