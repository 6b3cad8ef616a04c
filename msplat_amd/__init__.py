"""msplat_amd: MI355X-native 3D Gaussian Splatting training engine."""
