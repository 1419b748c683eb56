"""Synthetic column data for the benchmark workloads, generated on the GPU
from seeded torch generators (no files, no host round trip)."""
import torch

ZIPF_N = 10_000_000
ZIPF_S = 1.1


def multirun_keys(n, device, seed=1):
    """int64 keys: 50% uniform over [0, 2^63), 50% Zipf(s=1.1) ranks over
    10^7 distinct values [0, 10^7) (rank 0 most frequent), chosen per row by
    a fair coin."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    coin = torch.randint(0, 2, (n,), generator=g, device=device, dtype=torch.int32).bool()
    hi = torch.randint(0, 2**31, (n,), generator=g, device=device, dtype=torch.int64)
    lo = torch.randint(0, 2**32, (n,), generator=g, device=device, dtype=torch.int64)
    uni = (hi << 32) | lo
    del hi, lo
    ranks = torch.arange(1, ZIPF_N + 1, device=device, dtype=torch.float64)
    cdf = torch.cumsum(ranks.pow(-ZIPF_S), 0)
    cdf /= cdf[-1].clone()
    del ranks
    u = torch.rand(n, generator=g, device=device, dtype=torch.float64)
    zipf = torch.searchsorted(cdf, u).clamp_(max=ZIPF_N - 1)
    del u, cdf
    keys = torch.where(coin, uni, zipf)
    del uni, zipf, coin
    return keys.contiguous()
