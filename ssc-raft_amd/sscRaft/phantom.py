"""Seeded random-ellipse phantoms and noisy parallel-beam sinograms.

Synthetic inputs for the benchmarks and tests (the reference ships no
phantom generator).  A phantom lives on [-1, 1]^2 sampled at pixel centres
x_i = -1 + (i + 0.5) * 2/N, holds 10-20 ellipses with centres uniform in the
disk of radius 0.7, semi-axes ~ U[0.05, 0.5], rotation ~ U[0, pi) and values
~ U[0, 1]; the values are summed and the sum is clipped to [0, 1].  The RNG is
numpy's Generator(PCG64(seed)).

``ellipse_params`` draws the parameters on the host; ``rasterize`` evaluates
them either with numpy or, for large batches, with torch on a GPU device (the
same arithmetic in float32).
"""
import numpy as np


def ellipse_params(seed, min_ellipses=10, max_ellipses=20):
    """Return an (E, 6) float64 array of (cx, cy, a, b, phi, value)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ne = int(rng.integers(min_ellipses, max_ellipses + 1))
    r = 0.7 * np.sqrt(rng.random(ne))
    ang = 2.0 * np.pi * rng.random(ne)
    cx, cy = r * np.cos(ang), r * np.sin(ang)
    a = rng.uniform(0.05, 0.5, ne)
    b = rng.uniform(0.05, 0.5, ne)
    phi = rng.uniform(0.0, np.pi, ne)
    val = rng.random(ne)
    return np.stack([cx, cy, a, b, phi, val], axis=1)


def _grid(n):
    return (-1.0 + (np.arange(n) + 0.5) * (2.0 / n)).astype(np.float32)


def rasterize(params, n, device=None):
    """Rasterise one ellipse list onto an n x n float32 image [y][x]."""
    if device is None:
        g = _grid(n)
        x = g[None, :]
        y = g[:, None]
        img = np.zeros((n, n), dtype=np.float32)
        for cx, cy, a, b, phi, v in params:
            c, s = np.float32(np.cos(phi)), np.float32(np.sin(phi))
            xr = (x - np.float32(cx)) * c + (y - np.float32(cy)) * s
            yr = -(x - np.float32(cx)) * s + (y - np.float32(cy)) * c
            inside = (xr / np.float32(a)) ** 2 + (yr / np.float32(b)) ** 2 <= 1.0
            img += np.where(inside, np.float32(v), np.float32(0))
        return np.clip(img, 0.0, 1.0)
    import torch

    g = torch.as_tensor(_grid(n), device=device)
    x = g[None, :]
    y = g[:, None]
    img = torch.zeros((n, n), dtype=torch.float32, device=device)
    for cx, cy, a, b, phi, v in params:
        c, s = float(np.float32(np.cos(phi))), float(np.float32(np.sin(phi)))
        xr = (x - float(np.float32(cx))) * c + (y - float(np.float32(cy))) * s
        yr = -(x - float(np.float32(cx))) * s + (y - float(np.float32(cy))) * c
        inside = (xr / float(np.float32(a))) ** 2 + (yr / float(np.float32(b))) ** 2 <= 1.0
        img += inside.to(torch.float32) * float(np.float32(v))
    return img.clamp_(0.0, 1.0)


def phantoms(n, count, config=2, first=0, device=None):
    """Stack of `count` phantoms with seeds 1000*config + first + index."""
    if device is None:
        return np.stack([rasterize(ellipse_params(1000 * config + first + k), n) for k in range(count)])
    import torch

    return torch.stack([rasterize(ellipse_params(1000 * config + first + k), n, device) for k in range(count)])


def uniform_angles(nangles):
    """float32(linspace(0, pi, A, endpoint=False))."""
    return np.linspace(0.0, np.pi, nangles, endpoint=False).astype(np.float32)


def poisson_log(p, i0, seed):
    """counts ~ Poisson(I0 exp(-p)); returns p_hat = -ln(max(counts, 1)/I0) (numpy, float32)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = rng.poisson(i0 * np.exp(-np.asarray(p, dtype=np.float64)))
    return (-np.log(np.maximum(counts, 1) / i0)).astype(np.float32)


def poisson_log_torch(p, i0, seed):
    """Torch/GPU version of poisson_log (torch generator seeded with `seed`)."""
    import torch

    gen = torch.Generator(device=p.device)
    gen.manual_seed(int(seed))
    lam = i0 * torch.exp(-p.to(torch.float64))
    counts = torch.poisson(lam, generator=gen)
    return (-torch.log(torch.clamp(counts, min=1.0) / i0)).to(torch.float32)
