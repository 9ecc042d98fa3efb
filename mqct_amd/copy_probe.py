"""A plain device copy of a known size (256 MiB read, 256 MiB written, an
elementwise kernel) for scaling the HBM byte counters that rocprofv3 reports
per dispatch (FETCH_SIZE, WRITE_SIZE, in KiB):

    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex elementwise -- python3 -m mqct_amd.copy_probe

The calibration factor is 262144 KiB divided by the counter's reading.
"""
import sys


def main():
    import torch
    n = (256 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(4):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print("copy probe: 4 launches of %d bytes read + %d bytes written" % (4 * n, 4 * n))
    return 0


if __name__ == "__main__":
    sys.exit(main())
