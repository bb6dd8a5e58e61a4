"""pg_gpu_amd -- MI355X-native two-locus LD (DD / Dz / pi2, distance-binned).

A drop-in for pg_gpu's binned LD path: the containers
(HaplotypeMatrix, GenotypeMatrix), the per-pair estimator API
(ld_statistics.dd / dz / pi2 / r / r_squared / d_prime) and the binned
drivers (moments_ld.compute_ld_statistics,
HaplotypeMatrix.compute_ld_statistics_gpu_single_pop / _two_pops), running on
hand-written HIP kernels for gfx950 through the C ABI in include/pgld.h.
"""
__version__ = "0.1.0"
