"""Known-size device copy, the reference for rocprof's HBM byte counters.

    python -m spectralbte_amd.calibrate --gib 4 --reps 3

copies a buffer of the given size with the library's plain streaming kernel
(spbte_device_copy) `reps` times; FETCH_SIZE / WRITE_SIZE of that kernel per
dispatch, divided by the bytes moved, give the counters' scale factors."""
import argparse

import numpy as np

import spectralbte_amd as spb


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gib", type=float, default=4.0)
    p.add_argument("--reps", type=int, default=3)
    a = p.parse_args()
    n = int(a.gib * (1 << 30)) // 16 * 2
    src = spb.DeviceArray((n,), np.zeros(n))
    dst = spb.DeviceArray((n,))
    for _ in range(a.reps):
        spb._check(spb.lib().spbte_device_copy(dst.ptr, src.ptr, 8 * n))
    print(f"copied {8 * n} bytes x {a.reps}")


if __name__ == "__main__":
    main()
