"""Machine probes behind the roofline statements (libvflip_probe.so, vflip_amd/probe/hbm_probe.hip).

    python -m vflip_amd.bench_stream [--n 16777216] [--reps 20]

Prints one JSON line:
* streaming bandwidth of plain passes over arrays of n floats (256^3 by default,
  the level-0 size of the 256^3 solve), float4 per lane, for grids of 256..8192
  workgroups: a copy (8 B per element), the FP32 sweep's access pattern of six
  read streams and one write stream (28 B per element), a read-only sum (4 B), and
  the mixed-precision sweep's 20 B per element laid out as seven streams (four
  2-byte coefficient arrays) or as four (one 8-byte coefficient record per
  element); the best rate of each is the ceiling a one-pass kernel of that shape
  reaches;
* the cost of one grid-wide barrier of a cooperative launch against the number of
  workgroups, and of back-to-back one-word kernel launches.
"""
import argparse
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvflip_probe.so")


def load():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is missing: run vflip_amd.build.build() first")
    lib = ctypes.CDLL(LIB)
    lib.vflip_probe_stream.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.vflip_probe_grid_barrier.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double)]
    lib.vflip_probe_launch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
    return lib


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256 ** 3)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        print("bench_stream needs a GPU", file=sys.stderr)
        return 1
    lib = load()
    n = a.n - a.n % 4
    # bufs[2] and bufs[3] are adjacent halves of one allocation: the 8-byte records of kind 4
    pair = torch.rand(2 * n, device="cuda")
    bufs = [torch.rand(n, device="cuda"), torch.rand(n, device="cuda"), pair[:n], pair[n:]] + \
        [torch.rand(n, device="cuda") for _ in range(3)]
    ptrs = (ctypes.c_void_p * 7)(*[b.data_ptr() for b in bufs])
    kinds = {"copy": (0, 8), "six_reads_one_write": (1, 28), "read_sum": (2, 4),
             "sweep_fp16_seven_streams": (3, 20), "sweep_fp16_four_streams": (4, 20)}
    out = {"n": n, "array_MB": n * 4 / 1e6, "streams": {}}
    for name, (kind, bytes_per) in kinds.items():
        rows = []
        for blocks in (256, 1024, 2048, 4096, 8192):
            ms = ctypes.c_double()
            rc = lib.vflip_probe_stream(kind, ptrs, n, blocks, a.reps, ctypes.byref(ms))
            if rc:
                raise RuntimeError(f"vflip_probe_stream({name}, {blocks}) failed: {rc}")
            rows.append({"workgroups": blocks, "us": round(ms.value * 1e3, 2),
                         "GBps": round(bytes_per * n / (ms.value * 1e-3) / 1e9, 1)})
        best = max(rows, key=lambda r: r["GBps"])
        out["streams"][name] = {"bytes_per_element": bytes_per, "runs": rows, "best_GBps": best["GBps"],
                                "best_frac_of_8TBps": round(best["GBps"] / 8000.0, 3)}
    slot = torch.zeros(8192, device="cuda")
    bar = []
    for blocks in (8, 32, 64, 128, 256, 512):
        us = ctypes.c_double()
        rc = lib.vflip_probe_grid_barrier(ctypes.c_void_p(slot.data_ptr()), blocks, 200, ctypes.byref(us))
        bar.append({"workgroups": blocks, "us_per_barrier": round(us.value, 2) if rc == 0 else None, "rc": rc})
    out["grid_barrier"] = bar
    g, h = ctypes.c_double(), ctypes.c_double()
    rc = lib.vflip_probe_launch(ctypes.c_void_p(slot.data_ptr()), 1024, 200, ctypes.byref(g), ctypes.byref(h))
    out["small_launch"] = {"workgroups": 1024, "gpu_us_per_launch": round(g.value, 2),
                           "host_us_per_enqueue": round(h.value, 2), "rc": rc}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
