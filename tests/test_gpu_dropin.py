"""The drop-in symbols for PyProp's generated tensor-potential kernels
(include/pyprop_amd_dropin.h), called by a separate host program compiled against those C
signatures and linked against libpyprop_amd.so (tests/dropin/caller.cpp), with the
generated wrapper's argument list (generator.py:1428-1451): dest += scaling V source
against the oracle's restatement of the generated loops, 1e-12 relative, with a non-zero
dest to check accumulation.  Helium (Simp x Simp x BandNH, small basis) and the hydrogen
workload's two potentials (Diag x BandNH, Simp x BandNH)."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.conftest import crandn

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CALLER = ROOT / "tests" / "dropin" / "_build" / "caller"
KIND = {("Simp", "Simp", "BandNH"): 0, ("Diag", "BandNH"): 1, ("Simp", "BandNH"): 2}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(tmp_path, ids, V, scaling, src, dst, pairs):
    kind = KIND[tuple(ids)]
    rank = len(ids)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([kind, rank], np.int32).tofile(f)
        for a in (V, src, dst):
            np.array(a.shape, np.int32).tofile(f)
        np.array([scaling], np.float64).tofile(f)
        for a in (V, src, dst):
            np.ascontiguousarray(a, np.complex128).tofile(f)
        for s, p in zip(ids, pairs):
            if s == "Simp":
                np.array([len(p)], np.int32).tofile(f)
                np.ascontiguousarray(p, np.int32).tofile(f)
    out = tmp_path / "out.bin"
    r = subprocess.run([str(CALLER), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and not r.stderr, r.stderr
    return np.fromfile(out, np.complex128).reshape(dst.shape)


def _systems():
    from pyprop_amd.systems import helium_system, hydrogen_system
    he = helium_system(lmax=2, L=(0, 1, 2), order=5, n_breakpoints=12, rmax=15.0, n_inner=6, gamma=2.5)
    h = hydrogen_system()
    return [("helium " + p.name[:24], he.shape, p) for p in he.potentials] + \
           [("hydrogen " + p.name[:24], h.shape, p) for p in h.potentials]


@pytest.mark.parametrize("case", range(5))
def test_dropin_symbol_vs_oracle(tmp_path, case):
    assert CALLER.exists(), "build the drop-in caller first (__graft_entry__.build)"
    name, shape, p = _systems()[case]
    say("drop-in %s %s via the caller program" % ("_".join(p.storage_ids), name))
    rng = np.random.default_rng(40 + case)
    src = crandn(rng, shape)
    dst0 = crandn(rng, shape)
    pairs = [g.pairs if g.storage_id == "Simp" else None for g in p.geometries]
    got = _run(tmp_path, p.storage_ids, p.data, 0.37, src, dst0, pairs)
    ref = oracle.tpmul(p.storage_ids, p.data, 0.37, src, dst0.copy(), pairs)
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref - dst0)
    say("  |caller - oracle| / |V src| = %.2e" % e)
    assert e < 1e-12
