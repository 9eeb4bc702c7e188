"""fk_mc_amd -- MI355X-native per-move Hamiltonian diagonalisation path of fk_mc.

The product is the HIP library fk_mc_amd/lib/libfkgpu.so behind the C ABI in
include/fkgpu.h, plus the C++ host layer (fk_mc_amd/host) that keeps the
reference's configuration / move / Metropolis / measure interfaces.  This Python
package is a thin ctypes binding used by the tests and the benchmark.
"""
from .fkgpu import (Context, Chains, FkgpuError, LATTICE_HYPERCUBIC, LATTICE_TRIANGULAR,  # noqa: F401
                    MOVE_ADDREMOVE, MOVE_FLIP, MOVE_HOP, MOVE_RANDOMIZE, SOLVER_AUTO, SOLVER_VARIANTS, PRECISION_FP64, PRECISION_MIXED, UPDATE_DENSE, UPDATE_RANK1, KPM_DELTA_LOCAL, KPM_DELTA_STOCHASTIC,
                    MEASURE_SPECTRUM_HISTORY, MEASURE_FOCC_HISTORY, MEASURE_IPR, MEASURE_STIFFNESS,
                    MEASURE_EIGENFUNCTIONS, MEASURE_NO_SPECTRUM, lib)
