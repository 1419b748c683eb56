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


STAR_DIM_ROWS = (2556, 2000, 30000, 200000)  # date, supplier, customer, part


def star_schema(n, device, seed=2, attrs=False):
    """Star-join fact table and dimensions: four int32 foreign keys uniform
    over the dimension rows, extendedprice int32 in [90000, 10^7] cents,
    discount int32 in [0, 10]; date.year = 1992 + floor(7 i / 2556),
    customer.region in [0, 7) filtered = 0, part.category in [0, 25)
    filtered = 0, part.brand in [0, 1000).  Returns the dimension list of
    ops.star_join_group_sum (GROUP BY year x brand, 7,000 groups) and the two
    measures; with `attrs`, also {"region": customer.region, "cat":
    part.category} for filters with other constants."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    fks = [torch.randint(0, R, (n,), generator=g, device=device, dtype=torch.int32) for R in STAR_DIM_ROWS]
    m1 = torch.randint(90000, 10000001, (n,), generator=g, device=device, dtype=torch.int32)
    m2 = torch.randint(0, 11, (n,), generator=g, device=device, dtype=torch.int32)
    year = (1992 + (7 * torch.arange(2556, device=device)) // 2556).to(torch.int32)
    region = torch.randint(0, 7, (30000,), generator=g, device=device, dtype=torch.int32)
    cat = torch.randint(0, 25, (200000,), generator=g, device=device, dtype=torch.int32)
    brand = torch.randint(0, 1000, (200000,), generator=g, device=device, dtype=torch.int32)
    dims = [{"fk": fks[0], "rows": 2556, "group": year, "group_base": 1992, "group_card": 7},
            {"fk": fks[1], "rows": 2000},
            {"fk": fks[2], "rows": 30000, "pass": (region == 0).to(torch.uint8)},
            {"fk": fks[3], "rows": 200000, "pass": (cat == 0).to(torch.uint8), "group": brand, "group_card": 1000}]
    if attrs:
        return dims, m1, m2, {"region": region, "cat": cat}
    return dims, m1, m2
