"""Write tests/golden/fftpack_rfft.npz: inputs and outputs of the reference's own RFFTF and
RFFTB (src/fftpack.f:2151, 2561), compiled from its sources by `make -C oracle ref` into
oracle/_ref/libfftpack_ref.so. Run in a container that has /root/reference; the fixture is
data (seeded inputs and the reference's outputs) so the tests need no reference at run time.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

LENGTHS = (2, 4, 6, 8, 12, 15, 16, 30, 32, 48, 64, 96, 128, 192, 256, 384, 512)


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    rng = np.random.default_rng(20240601)
    out = {}
    for L in LENGTHS:
        x = rng.uniform(-1.0, 1.0, size=(3, L))
        out["x_%d" % L] = x
        out["rfftf_%d" % L] = oracle.ref_rfftf(x)
        out["rfftb_%d" % L] = oracle.ref_rfftb(x)
    np.savez_compressed(os.path.join(HERE, "fftpack_rfft.npz"), **out)
    print("wrote", len(LENGTHS), "lengths")


if __name__ == "__main__":
    main()
