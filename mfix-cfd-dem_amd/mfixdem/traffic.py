"""HBM traffic of the fused step kernel from rocprofv3 counters.

  run        build the 1M-sphere fluidised bed, step it eagerly (no graph, so
             every launch is a dispatch the counters see), then launch a
             plain device-to-device copy of known size for calibration.
  summarize  read two counter_collection.csv files (one FETCH_SIZE pass,
             one WRITE_SIZE pass), scale each counter by what it reads for
             the copy kernel (bytes moved / counter), and write the bytes
             per k_step launch to a JSON file that bench.py reports.

  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o fetch -- python3 -m mfixdem.traffic run
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT -o write -- python3 -m mfixdem.traffic run
  python3 -m mfixdem.traffic summarize OUT/fetch_counter_collection.csv \\
      OUT/write_counter_collection.csv profiles/traffic_k_step_fluidised_1M.json
"""
from __future__ import annotations

import csv
import json
import os
import sys
import time

COPY_BYTES = 1 << 30   # 1 GiB read + 1 GiB written by the calibration copy
STEP_KERNEL = "k_step<false,0,false,false,false>"   # FP64, fused, GARG, single system, <= 8 facets


KERNEL_SOURCES = ("csrc/kernels.hpp", "csrc/device_physics.hpp", "csrc/params.hpp")


def kernel_source_digest() -> str:
    """sha256 over the device sources of the step kernel: the counter figures
    describe the kernel built from exactly these sources."""
    import hashlib
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        h.update(open(os.path.join(here, rel), "rb").read())
    return h.hexdigest()[:16]


def run(steps=12):
    from . import abi, workloads
    from .system import DemSystem
    p, parts, gas = workloads.fluidised_bed(n=1_000_000)
    s = DemSystem(p)
    s.set_particles(parts)
    s.set_gas(gas)
    s.set_debug(abi.DEBUG_NO_GRAPH)
    s.init()
    s.run(1, 30)          # past the first rebuild and warm-up
    s.run(31, steps)
    s.synchronize()
    s.close()
    # calibration: a device-to-device copy through the same HIP runtime the
    # library links (the runtime runs it as its copyBuffer blit kernel)
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(COPY_BYTES)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print("traffic run done", flush=True)


def _per_dispatch(path, counter):
    """{kernel name: [KiB per dispatch]} for one counter."""
    out = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        out.setdefault(r["Kernel_Name"], {}).setdefault(r["Dispatch_Id"], 0.0)
        out[r["Kernel_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in out.items()}


def summarize(fetch_csv, write_csv, out_json):
    f = _per_dispatch(fetch_csv, "FETCH_SIZE")
    w = _per_dispatch(write_csv, "WRITE_SIZE")

    def pick(d, pred):
        ks = [k for k in d if pred(k)]
        if not ks:
            raise SystemExit(f"no kernel matches in {sorted(d)}")
        vals = [v for k in ks for v in d[k]]
        return sum(vals) / len(vals)

    is_step = lambda k: STEP_KERNEL in k.replace(" ", "")
    # the calibration copy: the largest dispatch of a copy / elementwise kernel
    def copy_kib(d):
        best = 0.0
        for k, vals in d.items():
            if "copy" in k.lower() or "elementwise" in k.lower():
                best = max(best, max(vals))
        return best
    fc, wc = copy_kib(f), copy_kib(w)
    scale_r = COPY_BYTES / (fc * 1024.0)
    scale_w = COPY_BYTES / (wc * 1024.0)
    fk, wk = pick(f, is_step), pick(w, is_step)
    res = {
        "kernel": "k_step (fused, FP64, GARG_2012), fluidised bed 1M spheres",
        "fetch_kib_per_launch": fk, "write_kib_per_launch": wk,
        "copy_fetch_kib": fc, "copy_write_kib": wc,
        "read_scale": scale_r, "write_scale": scale_w,
        "bytes_per_launch": fk * 1024.0 * scale_r + wk * 1024.0 * scale_w,
        "measured_utc": time.strftime("%Y-%m-%dT%H:%MZ", time.gmtime(os.path.getmtime(fetch_csv))),
        "kernel_source_sha256_16": kernel_source_digest(),
    }
    json.dump(res, open(out_json, "w"), indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    if sys.argv[1] == "run":
        run()
    else:
        summarize(*sys.argv[2:5])
