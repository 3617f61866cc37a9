"""Run the reference driver (oracle/_ref/ref_driver) and parse its binary output.

Only tests/golden/make_golden.py uses this, in the container where the reference tree
exists; the committed fixtures it writes are what the tests read.
"""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


class Reader:
    def __init__(self, data):
        self.b = data
        self.o = 0

    def f(self, n=None):
        k = 1 if n is None else int(n)
        v = np.frombuffer(self.b, dtype="<f8", count=k, offset=self.o).copy()
        self.o += 8 * k
        return float(v[0]) if n is None else v

    def i(self, n=None):
        k = 1 if n is None else int(n)
        v = np.frombuffer(self.b, dtype="<i8", count=k, offset=self.o).copy()
        self.o += 8 * k
        return int(v[0]) if n is None else v

    def done(self):
        return self.o == len(self.b)


def pack(*items):
    out = []
    for it in items:
        a = np.asarray(it)
        if a.dtype.kind in "iu":
            out.append(a.astype("<i8").tobytes())
        else:
            out.append(a.astype("<f8").tobytes())
    return b"".join(out)


def run(mode, deck, inbytes, timeout=3600, files=()):
    """Run the driver in a scratch directory; returns (outfile bytes, stdout) and, when files
    names text files the driver writes there, also their contents."""
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(inbytes)
        env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1")
        r = subprocess.run([DRIVER, mode, fin, fout], input=deck.encode(), cwd=td,
                           capture_output=True, timeout=timeout, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"ref_driver {mode} failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        with open(fout, "rb") as fh:
            data = fh.read()
        if files:
            texts = {}
            for f in files:
                with open(os.path.join(td, f)) as fh:
                    texts[f] = fh.read()
            return data, r.stdout.decode(errors="replace"), texts
        return data, r.stdout.decode(errors="replace")
