"""Writes exact_leverages.npz: the reference's akm_kss_decompose with exact
leverages (LeverageMethod::Exact, match level, no pruning) on the JLA golden
panel, through the reference CPU core compiled under oracle/_ref (checker
only): its working rows, P_ii, variance components (plug-in, AGSU, KSS;
sigma2_ho, n_py) and sigma_i.

    python tests/golden/make_exact_leverages.py      (needs oracle/_ref)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path[:0] = [ROOT, os.path.join(ROOT, "xhdfe-xfe_amd"), HERE]

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from golden_cases import jla_panel  # noqa: E402


def main():
    jp = jla_panel()
    rw, rf, rc, pii, comp, sig = oracle.ref_akm_exact(jp["y"], jp["worker"], jp["firm"])
    path = os.path.join(HERE, "exact_leverages.npz")
    np.savez_compressed(path, row_worker=rw, row_firm=rf, row_weight=rc, pii=pii, components=comp, sigma_i=sig)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
