"""The pair loops of the HIP library divide by per-particle quantities (h, h^3, h^4, rho Omega) and
by |r_ab| through precomputed reciprocals: q = RN(x r), e = x - q y (exact, one FMA),
x / y = RN(q + e r) with r = RN(1 / y) (Markstein, IBM J. Res. Dev. 34(1), 1990).  That identity
is what keeps those loops bit-identical to the reference's IEEE divisions; this test checks it
with the host's IEEE double arithmetic and C99 fma on 10^7 random and adversarial pairs."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fma_corrected_division_is_correctly_rounded(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "div_check")
    subprocess.run([cc, "-O2", "-ffp-contract=off", "-std=c99",
                    os.path.join(HERE, "division_identity_check.c"), "-o", exe, "-lm"],
                   check=True, capture_output=True)
    out = subprocess.run([exe, "10000000"], check=True, capture_output=True, text=True).stdout
    assert out.strip().splitlines()[-1] == "0", out
