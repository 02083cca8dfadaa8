"""bench.py's issued-MFMA accounting of the factored multiply against the instruction
counts the hardware reports for the same kernels at helium (SQ_INSTS_MFMA per dispatch,
rocprofv3 --pmc on MI355X): band-tile radial kernels 1,487,808, position-form radial
kernels 928,800 (with and without the window in LDS), angular kernels 259,200 in every GEMM
form and none in the sparse forms (vector FMAs).  The one-term band-tile kernel ("1x1") issues padded band batches beyond the
count here and is not covered."""
import pytest

import bench
from pyprop_amd.factored import FactoredHamiltonian


class _Helium:
    """The helium operator's shape as bench.factored_flops reads it: 18 terms, band
    half-width 6, 12 of the terms on the band tiles (the rest separable)."""
    nterms, bw, band_terms = 18, 6, 12
    angular_nnz = 1269  # non-zeros of the 18 angular matrices (42 channels, l <= 5, L <= 2)
    RADIAL_SHAPES = FactoredHamiltonian.RADIAL_SHAPES
    ANGULAR_VARIANTS = FactoredHamiltonian.ANGULAR_VARIANTS

    def __init__(self, radial, angular):
        self._r = self.RADIAL_SHAPES.index(radial)
        self._a = self.ANGULAR_VARIANTS.index(angular)

    def radial_variant(self, precision):
        return self._r

    def angular_variant(self, precision):
        return self._a


@pytest.mark.parametrize("angular", FactoredHamiltonian.ANGULAR_VARIANTS)
@pytest.mark.parametrize("radial, counted", [("2x2", 1487808), ("3x2", 1487808), ("3x3p", 1487808),
                                             ("pos6", 928800), ("pos3", 928800), ("pos2", 928800),
                                             ("pos3w16", 928800), ("pos2w16", 928800), ("pos2w8", 928800),
                                             ("pos1w16", 928800), ("pos2w16s", 928800)])
def test_issued_mfma_matches_hardware_counts(radial, counted, angular):
    issued, useful = bench.factored_flops(_Helium(radial, angular), 42, 60, 64)
    assert issued == (counted + (0 if angular.startswith("sparse") else 259200)) * 2048
    assert 0 < useful < issued
