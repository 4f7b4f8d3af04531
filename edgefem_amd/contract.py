"""The array-package artifact contract (python/edgefem/contract.py): a canonical 2x2 array
package with hand-pinned values, written through the array-package writers so that
consumers of the ``EdgeFEM_ArrayPackage`` JSON and the pattern and scan CSVs can pin their
loaders against it.  The fixture values are the contract's data (the format, not solver
output), so they are the reference's: 15 mm element spacing, 10 and 10.5 GHz, -15 dB
return at +30 deg on the diagonal, -20 dB at -45 deg between neighbours, -25 dB at +10 deg
across the diagonal (x 0.95 at the second frequency), cosine patterns with a port-dependent
phase ramp, and the hand-checkable scan reflection coefficients.

    python -m edgefem_amd.contract out_dir/

writes golden_array_package.json, golden_patterns.csv and golden_scan.csv.
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

from . import pyedgefem as em

FIXTURE_REVISION = 2

_NX, _NY = 2, 2
_SPACING_M = 0.015
_FREQS_HZ = (10.0e9, 10.5e9)
_THETA_DEG = np.arange(-90.0, 91.0, 5.0)
_SCAN_GAMMAS = {
    (0.0, 0.0): [0.10 + 0.00j, 0.12 + 0.02j, 0.12 - 0.02j, 0.10 + 0.00j],
    (30.0, 0.0): [0.20 + 0.10j, 0.25 + 0.05j, 0.25 - 0.05j, 0.20 - 0.10j],
    (60.0, 0.0): [0.50 + 0.00j, 0.30 - 0.40j, 0.30 + 0.40j, 0.50 + 0.00j],
}
_SCAN_LOSS_DB = {(0.0, 0.0): 0.0, (30.0, 0.0): 0.62, (60.0, 0.0): 3.01}


def _positions():
    return [np.array([ix * _SPACING_M, iy * _SPACING_M, 0.0]) for ix in range(_NX) for iy in range(_NY)]


def canonical_array_package() -> em.ArrayPackage:
    pkg = em.ArrayPackage()
    pkg.model_name = f"golden-2x2-rev{FIXTURE_REVISION}"
    pkg.reference_impedance = 50.0
    pkg.element_positions = _positions()
    pkg.frequencies = list(_FREQS_HZ)
    n = _NX * _NY
    S = np.empty((n, n), np.complex128)
    grid = [divmod(i, _NY) for i in range(n)]
    for i in range(n):
        for j in range(n):
            d = math.hypot(grid[i][0] - grid[j][0], grid[i][1] - grid[j][1])
            db, ang = (-15.0, 30.0) if i == j else ((-20.0, -45.0) if d == 1.0 else (-25.0, 10.0))
            S[i, j] = 10 ** (db / 20) * np.exp(1j * math.radians(ang))
    pkg.S_matrices = [S, S * 0.95]
    return pkg


def canonical_patterns():
    pats = []
    for port, pos in enumerate(_positions()):
        pat = em.EmbeddedPattern()
        pat.port_index = port
        pat.frequency = _FREQS_HZ[0]
        pat.element_position = pos
        for td in _THETA_DEG:
            th = math.radians(float(td))
            ph = math.radians(2.0 * port + 0.1 * float(td))
            pat.pattern_phi0.append(em.FFPoint2D(th, max(math.cos(th), 0.0), ph))
            pat.pattern_phi90.append(em.FFPoint2D(th, 0.9 * max(math.cos(th), 0.0), ph))
        pats.append(pat)
    return pats


def canonical_scan_results():
    z0 = canonical_array_package().reference_impedance
    out = []
    for (td, pd), gammas in _SCAN_GAMMAS.items():
        r = em.ActiveImpedanceResult()
        r.theta, r.phi = math.radians(td), math.radians(pd)
        g = np.array(gammas, np.complex128)
        r.Gamma_active = em.VecC(g)
        r.Z_active = em.VecC(z0 * (1 + g) / (1 - g))
        r.scan_loss_db = _SCAN_LOSS_DB[(td, pd)]
        out.append(r)
    return out


def write_golden_array_package(out_dir):
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    paths = (out / "golden_array_package.json", out / "golden_patterns.csv", out / "golden_scan.csv")
    em.export_array_package_json(str(paths[0]), canonical_array_package())
    em.export_patterns_csv(str(paths[1]), canonical_patterns())
    em.export_scan_csv(str(paths[2]), canonical_scan_results())
    return paths


if __name__ == "__main__":
    for p in write_golden_array_package(sys.argv[1] if len(sys.argv) > 1 else "."):
        print(f"wrote {p}")
