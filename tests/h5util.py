"""Minimal HDF5 file access for the tests through the h5import / h5dump tools
(no Python HDF5 binding is installed in this image)."""
import os
import shutil
import subprocess
import tempfile

import numpy as np


def tool(name):
    for cand in (shutil.which(name), "/opt/conda/bin/" + name):
        if cand and os.path.exists(cand):
            return cand
    return None


def available():
    return tool("h5import") is not None and tool("h5dump") is not None


def write_matrix(path, name, A):
    """Create (or extend) `path` with a 2-D float64 dataset `name`."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    with tempfile.TemporaryDirectory() as tmp:
        txt = os.path.join(tmp, "a.txt")
        cfg = os.path.join(tmp, "a.cfg")
        np.savetxt(txt, A.reshape(A.shape[0], -1), fmt="%.17g")
        dims = " ".join(str(d) for d in A.shape)
        with open(cfg, "w") as f:
            f.write(f"PATH {name}\nINPUT-CLASS TEXTFP\nRANK {A.ndim}\nDIMENSION-SIZES {dims}\n"
                    "OUTPUT-CLASS FP\nOUTPUT-SIZE 64\nOUTPUT-ARCHITECTURE IEEE\n"
                    "OUTPUT-BYTE-ORDER LE\n")
        subprocess.run([tool("h5import"), txt, "-c", cfg, "-o", path], check=True,
                       stdout=subprocess.DEVNULL)


def read(path, name, dtype="<f8"):
    """Dataset `name` of `path` as a flat numpy array (raw little-endian export)."""
    with tempfile.TemporaryDirectory() as tmp:
        b = os.path.join(tmp, "d.bin")
        subprocess.run([tool("h5dump"), "-d", "/" + name, "-b", "LE", "-o", b, path], check=True,
                       stdout=subprocess.DEVNULL)
        return np.fromfile(b, dtype=dtype)


def shape(path, name):
    out = subprocess.run([tool("h5dump"), "-H", "-d", "/" + name, path], check=True,
                         capture_output=True, text=True).stdout
    line = [l for l in out.splitlines() if "DATASPACE" in l][0]
    inner = line.split("(")[1].split(")")[0]
    return tuple(int(x) for x in inner.split(","))
