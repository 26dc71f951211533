"""Every entry point with gsl_matrix-style padded rows (tda > columns,
comm.c:1866-1883), called from C the way INTEGRATION.md's swap calls it: the
tests/c/tda_caller program runs each problem with dense rows and with
64-byte-padded rows holding NaN in the padding, and checks that all outputs
agree bit for bit and that no output padding is written."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "c", "tda_caller")


@pytest.mark.parametrize("prec,N,D,K,M1,binary", [
    (0, 3000, 166, 343, 28, 1),     # FP64, fingerprint width: tda 168, M+1 tda 32
    (1, 3000, 166, 343, 28, 1),     # mixed
    (2, 3000, 166, 343, 28, 1),     # mixed, single-term R
    (0, 2000, 12, 100, 17, 0),      # FP64, float data: tda 16, M+1 tda 24
    (1, 2000, 12, 100, 17, 0),      # mixed, float data (hi/lo data terms)
    (1, 5000, 64, 216, 126, 1),     # D = 64 (dense rows already aligned: tda 72 forced)
])
def test_padded_rows_match_dense(prec, N, D, K, M1, binary):
    if not os.path.exists(CALLER):
        pytest.fail("tests/c/tda_caller is not built (make -C tests/c)")
    print(f"case: tda caller precision {prec} N={N} D={D} K={K} M1={M1}", flush=True)
    sys.stdout.flush()
    r = subprocess.run([CALLER, str(prec), str(N), str(D), str(K), str(M1), str(binary), "7"],
                       capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "tda OK" in r.stdout, r.stdout + r.stderr
