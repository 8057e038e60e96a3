"""pgmhip: the MI355X-native pGM electrostatics engine (host side).

The engine itself is the HIP library csrc/ builds (libpgm_hip.so, C ABI in include/pgm_hip.h); this package
holds its Python binding (`engine`), the prmtop reader (`prmtop`) and the seeded workload generators
(`systems`).
"""
