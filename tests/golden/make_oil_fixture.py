"""Regenerate ``oil_state.npz`` / ``oil_input.npz`` from the reference's HDF5 data files.

The reference ships two HDF5 files (``data/Oil.h5``: the input T/id/lb of the
"Oil flow" set; ``data/Oil.out.h5``: a full dagtm output state at temperature 1,
K=125, L=3, M=8+bias, N=1000, D=12).  No Python HDF5 binding is installed, so
each dataset is exported with ``h5dump -b LE`` (raw little-endian bytes, no
text round trip) and re-read with numpy.  Only data is kept: the arrays are the
reference's own output vectors, used as golden values by the tests.

Run in a container that has ``/root/reference`` and an ``h5dump`` binary:

    python tests/golden/make_oil_fixture.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference/data"
HERE = os.path.dirname(os.path.abspath(__file__))

# (dataset, dtype, shape) as listed by `h5dump -H`
OUT_SETS = [
    ("FI.dat", "<f8", (125, 9)),
    ("M.dat", "<f8", (1000, 3)),
    ("R.dat", "<f8", (125, 1000)),
    ("T.dat", "<f8", (1000, 12)),
    ("W.dat", "<f8", (9, 12)),
    ("X.dat", "<f8", (125, 3)),
    ("Y.dat", "<f8", (125, 12)),
    ("beta.dat", "<f8", (1,)),
    ("lb.dat", "<i4", (1000,)),
]
IN_SETS = [
    ("T", "<f4", (1000, 12)),
    ("id", "<i4", (1000,)),
    ("lb", "<i4", (1000,)),
]


def _h5dump():
    for cand in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if cand and os.path.exists(cand):
            return cand
    sys.exit("h5dump not found")


def _export(h5file, sets, tmp):
    out = {}
    for name, dt, shape in sets:
        binf = os.path.join(tmp, name + ".bin")
        subprocess.run([_h5dump(), "-d", "/" + name, "-b", "LE", "-o", binf, h5file],
                       check=True, stdout=subprocess.DEVNULL)
        arr = np.fromfile(binf, dtype=dt)
        assert arr.size == int(np.prod(shape)), (name, arr.size, shape)
        out[name.replace(".dat", "")] = arr.reshape(shape)
    return out


def main():
    with tempfile.TemporaryDirectory() as tmp:
        state = _export(os.path.join(REF, "Oil.out.h5"), OUT_SETS, tmp)
        inp = _export(os.path.join(REF, "Oil.h5"), IN_SETS, tmp)
    np.savez_compressed(os.path.join(HERE, "oil_state.npz"), **state)
    np.savez_compressed(os.path.join(HERE, "oil_input.npz"), **inp)
    print("wrote", sorted(state), sorted(inp))


if __name__ == "__main__":
    main()
