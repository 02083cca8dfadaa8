/* pyprop_amd drop-in symbols for PyProp's generated tensor-potential kernels.
 *
 * The same names and C signatures as the Fortran subroutines that
 * core/tensorpotential/tensorpotentialmultiply_generator.py writes for these storage
 * combinations (GetParameterList, generator.py:1373-1385; C declaration, GetWrapperCode,
 * generator.py:1428-1451; a concrete caller: examples/tensor/helium_stabilization/
 * benchmark/benchmark.cpp:11-14), so code compiled against the generated kernels links
 * against libpyprop_amd.so instead.
 *
 * All arguments are pointers; arrays are host memory (as blitz passes them), each followed
 * by its extents, extent 0 the slowest rank; Simp pair arrays are int32 [count][2]
 * (row = dest index, col = source index).  dest += scaling * V * source.  Each call stages
 * the arrays in HBM and copies dest back (the generated kernels' host-memory contract);
 * for HBM-resident products use pp_tp_create / pp_tp_multiply (pyprop_amd.h).  Like the
 * generated kernels these return nothing; a failure is reported on stderr and leaves dest
 * unchanged.
 */
#ifndef PYPROP_AMD_DROPIN_H
#define PYPROP_AMD_DROPIN_H

#include <complex>

extern "C" {

/* helium: coupled channel (Simp) x banded-packed r1 (Simp) x banded-nonhermitian r2 (BandNH) */
void tensorpotentialmultiply_simp_simp_bandnh_(std::complex<double>* potential, int* potentialExtent0,
                                               int* potentialExtent1, int* potentialExtent2, double* scaling,
                                               std::complex<double>* source, int* sourceExtent0, int* sourceExtent1,
                                               int* sourceExtent2, std::complex<double>* dest, int* destExtent0,
                                               int* destExtent1, int* destExtent2, int* pair0, int* pair0Extent0,
                                               int* pair0Extent1, int* pair1, int* pair1Extent0, int* pair1Extent1);

/* hydrogen H0: reduced spherical diagonal (Diag) x banded-nonhermitian r (BandNH) */
void tensorpotentialmultiply_diag_bandnh_(std::complex<double>* potential, int* potentialExtent0,
                                          int* potentialExtent1, double* scaling, std::complex<double>* source,
                                          int* sourceExtent0, int* sourceExtent1, std::complex<double>* dest,
                                          int* destExtent0, int* destExtent1);

/* hydrogen dipole: dipoleselectionrule (Simp) x banded-nonhermitian r (BandNH) */
void tensorpotentialmultiply_simp_bandnh_(std::complex<double>* potential, int* potentialExtent0,
                                          int* potentialExtent1, double* scaling, std::complex<double>* source,
                                          int* sourceExtent0, int* sourceExtent1, std::complex<double>* dest,
                                          int* destExtent0, int* destExtent1, int* pair0, int* pair0Extent0,
                                          int* pair0Extent1);
}

#endif
