"""Plain-copy calibration for rocprofv3's HBM traffic counters.

Copies a float32 buffer device-to-device (one copy kernel per call, N bytes
read and N bytes written) so that the FETCH_SIZE / WRITE_SIZE a counter pass
reports per dispatch can be compared with a known byte count:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o calib -- python3 -m chaste_amd.hbm_calib 24 1024

Sizes are in MiB; 24 MiB fits in the MI355X's 256 MiB Infinity Cache, 1024 MiB
does not.
"""
import sys


def main(argv):
    import torch
    sizes = [int(x) for x in argv] or [24, 1024]
    for mib in sizes:
        n = mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print('copied %d MiB x3' % mib, flush=True)
        del src, dst


if __name__ == '__main__':
    main(sys.argv[1:])
