"""Host-side physics coefficients (stopping power, scattering, removal)."""
