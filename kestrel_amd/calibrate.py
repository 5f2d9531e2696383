"""Device calibration: achievable FP64 and FP32 (non-packed, as this library
compiles FP32) vector FMA throughput, LDS read bandwidth and a plain HBM copy
of known size (the reference for hardware byte counters).

    python3 -m kestrel_amd.calibrate [--copy-bytes N]
"""
import argparse
import ctypes
import json

from . import _lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-bytes", type=int, default=256 << 20)
    a = ap.parse_args()
    L = _lib.load()
    tf, tf32, gb, lds = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _lib.check(L.kestrel_measure_fp64_peak(ctypes.byref(tf)))
    _lib.check(L.kestrel_measure_fp32_peak(ctypes.byref(tf32)))
    _lib.check(L.kestrel_measure_lds(ctypes.byref(lds)))
    _lib.check(L.kestrel_measure_copy(a.copy_bytes, ctypes.byref(gb)))
    print(json.dumps({"fp64_fma_tflops": tf.value, "fp32_fma_tflops": tf32.value, "lds_read_tbs": lds.value,
                      "copy_bytes": a.copy_bytes, "copy_gbps": gb.value}))


if __name__ == "__main__":
    main()
