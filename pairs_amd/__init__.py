"""MI355X-native particle hot path (cells, neighbour lists, ghosts, LJ/DEM
forces, velocity Verlet, thermo) behind the C ABI of include/pairs_hip.h."""
__version__ = "0.1.0"
