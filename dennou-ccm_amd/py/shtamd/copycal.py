"""Plain device copy of a known size (64 MiB, three dispatches).

Run under `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`: the counter
values it reads per dispatch against the 65536 KiB actually moved give the
factors that turn those counters into HBM bytes for the transform kernels.
"""
import torch

if __name__ == "__main__":
    n = 64 * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied", n * 8, "bytes x3", flush=True)
