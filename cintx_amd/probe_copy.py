"""Run the library's plain copy kernel on a known number of bytes.

Used under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to calibrate the byte
counters: `python3 -m cintx_amd.probe_copy [nelem] [reps]`.
"""
import ctypes
import sys

from . import capi


def main():
    nelem = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 26
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    L = capi.lib()
    L.cintx_copy_probe.argtypes = [ctypes.c_int64, ctypes.c_int, capi.DP]
    ms = ctypes.c_double()
    capi._check(L.cintx_copy_probe(nelem, reps, ctypes.byref(ms)))
    mb = nelem * 8 / 1e6
    print(f"copy {nelem} doubles ({mb:.1f} MB read + {mb:.1f} MB written): {ms.value:.3f} ms, "
          f"{2 * nelem * 8 / ms.value / 1e6:.1f} GB/s", flush=True)


if __name__ == "__main__":
    main()
