"""Synthetic workloads of the benchmarks, generated on the GPU from a seed.

uneven_batch: the batch of the grouped-DGEMM benchmark - count DGEMMs N/N with
m, n, k log-uniform in [64, 4096] (rounded), even entries U[-1,1), odd entries
(1 + U[0,1)) * (+-1) * 2^e with e uniform in [-12, 12]; alpha = 1, beta = 0."""
import math


def uneven_batch(torch, count, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    probs = []
    for i in range(count):
        m, n, k = (int(round(math.exp(rng.uniform(math.log(64), math.log(4096))))) for _ in range(3))
        wide = i % 2 == 1

        def gen(cnt):
            if not wide:
                return torch.rand(cnt, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
            e = torch.randint(-12, 13, (cnt,), device="cuda", generator=g).double()
            sg = torch.randint(0, 2, (cnt,), device="cuda", generator=g).double() * 2 - 1
            return (1 + torch.rand(cnt, dtype=torch.float64, device="cuda", generator=g)) * sg * torch.exp2(e)
        probs.append(dict(ta="N", tb="N", m=m, n=n, k=k, alpha=1.0, beta=0.0, A=gen(m * k), lda=m, B=gen(k * n),
                          ldb=k, C=torch.empty(m * n, dtype=torch.float64, device="cuda"), ldc=m, wide=wide))
    return probs
