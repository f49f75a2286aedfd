"""HBM calibration run: copies `--mib` MiB device to device with the library's
plain streaming copy kernel (k_copy), `--reps` times.  Profiled under
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, its per-dispatch readings against the
known byte count give the scale factor applied to the render kernel's
counters; timed with HIP events it gives the achievable copy bandwidth."""
import argparse
import ctypes as C
import importlib
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
K = importlib.import_module("3dgrt_kdtree_amd")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    dev = K.Device(0)
    n = a.mib * 1024 * 1024
    src, dst = dev.alloc(n), dev.alloc(n)
    K.check(K.lib().kdgrt_memset(dev.h, src.ptr, C.c_int(1), C.c_size_t(n)))
    s = dev.stream()
    K.check(K.lib().kdgrt_copy_f32x4(dev.h, dst.ptr, src.ptr, C.c_uint64(n // 16), s))
    dev.sync()
    dev.timer_begin(s)
    for _ in range(a.reps):
        K.check(K.lib().kdgrt_copy_f32x4(dev.h, dst.ptr, src.ptr, C.c_uint64(n // 16), s))
    ms = dev.timer_end(s) / a.reps
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "ms": round(ms, 4),
                      "copy_gbs": round(2 * n / (ms * 1e-3) / 1e9, 1)}))
    src.free()
    dst.free()
    dev.close()


if __name__ == "__main__":
    main()
