"""A plain device copy of known size, the yardstick for HBM traffic counters:
profiled under the same counters as a render, its FETCH_SIZE / WRITE_SIZE per
dispatch against the bytes it must move give the scale factor of the counters.

    python3 -m snelly_amd.copy_probe [MiB]     (default 1024 MiB, 5 copies)
"""
import sys


def main():
    import torch
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 5: each dispatch reads {mib * 1024} KiB and writes {mib * 1024} KiB", flush=True)


if __name__ == "__main__":
    main()
