"""raptor_amd: MI355X-native per-pixel geodesic + radiative-transfer path of RAPTOR."""
