"""Host packing of batched atoms (pyprop_amd.batch): the DIA arrays and couplings the
device kernel multiplies reproduce the operator the tensor potentials define (dense
SetupBigMatrix-style assembly, tests/test_oracle.py), and the atom sizes follow the
requested B-spline counts."""
import numpy as np

from pyprop_amd.batch import AtomSpec, pack_atom, random_specs
from pyprop_amd.systems import LaserSimpleLength
from tests.test_oracle import dense_operator


def dense_from_pack(h0, r, cpl, E):
    nl, w, n = h0.shape
    b = (w - 1) // 2
    H = np.zeros((nl * n, nl * n))
    for l in range(nl):
        for d in range(w):
            for i in range(n):
                j = i + d - b
                if 0 <= j < n:
                    H[l * n + i, l * n + j] += h0[l, d, i]
                    for q, lp in ((0, l - 1), (1, l + 1)):
                        if 0 <= lp < nl:
                            H[l * n + i, lp * n + j] += E * cpl[l, q] * r[d, i]
    return H


def test_pack_matches_dense_operator():
    spec = AtomSpec(3, 30, LaserSimpleLength(0.05, 0.1, 2.0), rmax=30.0, n_inner=8)
    sysm, h0, r, cpl, su, psi = pack_atom(spec)
    assert sysm.shape == (4, 30)
    E = 0.37
    H, S = dense_operator(sysm, [1.0, E])
    Hp = dense_from_pack(h0, r, cpl, E)
    assert np.abs(H - Hp).max() < 1e-12 * np.abs(H).max()
    # the initial state is S-normalised and lives in l = 0
    assert abs((psi.conj().ravel() @ S @ psi.ravel()).real - 1) < 1e-12
    assert np.all(psi[1:] == 0)


def test_random_specs_ranges():
    sp = random_specs(64, seed=5)
    assert all(80 <= s.n_bsplines <= 200 and 4 <= s.lmax <= 15 for s in sp)
    assert all(0.01 <= s.pulse.amplitude <= 0.1 and 0.05 <= s.pulse.frequency <= 0.3 for s in sp)
    sysm = pack_atom(sp[0])[0]
    assert sysm.shape == (sp[0].lmax + 1, sp[0].n_bsplines)
