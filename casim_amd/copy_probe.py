"""Plain device copy of a known size, for calibrating HBM byte counters.

Each dispatch reads 1 GiB and writes 1 GiB (torch's elementwise copy kernel);
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this program gives the factor
between counter readings and true HBM bytes.
"""
import torch

N = 256 * 1024 * 1024  # float32 elements = 1 GiB


def main():
    a = torch.ones(N, device="cuda")
    b = torch.empty_like(a)
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy probe: %d bytes read and written per dispatch" % (N * 4), flush=True)


if __name__ == "__main__":
    main()
