"""Device-to-device copy of a known size, used to calibrate the HBM byte
counters (TCC FETCH_SIZE / WRITE_SIZE) read by rocprofv3: the copy moves
exactly `bytes` in each direction, so counter / bytes is the scale factor
applied to the pair kernels' counter readings.

    python -m cvortex_amd.copy_probe [--bytes N]
"""
import argparse
import ctypes
import sys

from .hip import D2D, DeviceArray, check, hip, synchronize


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=1 << 29)
    args = ap.parse_args()
    n = args.bytes
    src = DeviceArray((n,), "u1")
    dst = DeviceArray((n,), "u1")
    src.fill_zero()
    dst.fill_zero()
    synchronize()
    check(hip().hipMemcpy(dst.ptr, src.ptr, ctypes.c_size_t(n), D2D), "hipMemcpy D2D")
    synchronize()
    print(f"copied {n} bytes device to device", file=sys.stderr)


if __name__ == "__main__":
    main()
