"""Reference copy for HBM counter calibration: one device-to-device copy of
1 GiB (torch), so that FETCH_SIZE / WRITE_SIZE of a kernel can be scaled by
what the same counters read for 1 GiB read + 1 GiB written."""
import torch

n = (1 << 30) // 8
a = torch.ones(n, dtype=torch.float64, device="cuda")
b = torch.empty_like(a)
for _ in range(3):
    b.copy_(a)
torch.cuda.synchronize()
print("copied 1 GiB x3", flush=True)
