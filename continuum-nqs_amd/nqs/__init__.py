"""MI355X-native VMC hot path of the continuum spinful Psiformer (drop-in for src/nqs of
wenqin-chen/continuum-nqs): psiformer, local_energy, ewald, sampler, minsr on hand-written HIP
kernels for gfx950 (libnqs_hip.so, next to this file)."""
