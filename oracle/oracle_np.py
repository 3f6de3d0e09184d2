"""Second, independent FP64 restatement of the same discretisation in NumPy.

TEST INFRASTRUCTURE ONLY (imported by tests/ as a cross-check of the C
oracle in raft_oracle.c).  It is written from the reference text, not from the
C oracle: the filter goes through numpy's pocketfft (rfft / irfft), and the
bin coordinates are evaluated the way the reference writes them (no fused
multiply-add), so the two oracles may disagree only for samples that sit
within ~1e-12 of a bin edge.

Reference (paths relative to the reference checkout):
  pad            cuda/src/common/padding.cu:140-178
  filter         cuda/src/processing/filters/filters.cu:33-51,163-190,395-438
  backprojection cuda/src/geometries/parallel/reconstruction_methods/fbp.cu:8-76
  Radon          cuda/src/geometries/parallel/forward_methods/radon_rt.cu:131-174
  eEM            cuda/src/geometries/parallel/reconstruction_methods/kernels/em_RT_kernel.cu:188-350,
                 reconstruction_methods/em_rt.cu:316-355
"""
import numpy as np

PLANK = 4.135667696e-15
VC = 299792458.0


def angle_steps(angles):
    a = np.asarray(angles, dtype=np.float32).astype(np.float64)
    d = np.empty_like(a)
    d[:-1] = a[1:] - a[:-1]
    d[-1] = a[-1] - a[-2]
    return d


def filter_apply(ftype, w, reg, pag):
    w = np.asarray(w, dtype=np.float64)
    pg = 1.0 / (1.0 + pag * w * w)
    if ftype == 1:
        return w * np.exp(-float(np.float32(0.693)) * reg * w * w) * pg
    if ftype == 2:
        return w * (1.0 / (1.0 + reg * w * w)) * pg
    if ftype == 3:
        return w * np.cos(np.pi * 0.5 * w) * pg
    if ftype == 4:
        param = np.maximum(w * reg * np.pi * 0.5, float(np.float32(1e-4)))
        return w * (np.sin(param) / param) * pg
    if ftype == 5:
        return w * (0.5 + 0.5 * np.cos(2 * np.pi * w)) * pg
    if ftype == 6:
        return w * (float(np.float32(0.54)) + float(np.float32(0.46)) * np.cos(2 * np.pi * w)) * pg
    if ftype == 7:
        return w * pg
    return np.ones_like(w)


def decode_fbp(paramf, parami):
    paramf = np.asarray(paramf, dtype=np.float32).astype(np.float64)
    nrays, nangles, nslices, objsize, pad = (int(v) for v in parami[:5])
    ftype = int(parami[7])
    bd, reg, energy, z2, px = paramf[0], paramf[1], paramf[2], paramf[3], paramf[4]
    offset = paramf[6]
    lam = 1.0 if bd == 0.0 else PLANK * VC / energy
    pag = lam * z2 * np.pi * bd / (px * px)
    return dict(nrays=nrays, nangles=nangles, nslices=nslices, pad=pad, ftype=ftype, reg=reg, pag=pag,
                offset=offset, pixel=px, npad=nrays * (1 + pad), L=(pad * nrays) // 2)


def pad_rows(tomo, npad, L):
    N = tomo.shape[-1]
    idx = np.clip(np.arange(npad) - L, 0, N - 1)
    return tomo[..., idx].astype(np.float64)


def filter_rows(tomo, paramf, parami):
    c = decode_fbp(paramf, parami)
    u = pad_rows(np.asarray(tomo, dtype=np.float32), c["npad"], c["L"])
    if c["ftype"] == 0:
        return u
    n = c["npad"]
    M = n // 2 + 1
    k = np.arange(M)
    w = 0.5 * np.minimum(k, M - k) / M
    H = filter_apply(c["ftype"], w, c["reg"], c["pag"]) * np.exp(-2j * np.pi * c["offset"] * k / (2 * M - 2))
    U = np.fft.rfft(u, axis=-1)
    # irfft = (1/n) * unnormalised C2R, so multiply back by n before /(n * pixel)
    q = np.fft.irfft(U * H, n=n, axis=-1) * n / (n * c["pixel"])
    return q


def fbp(tomo, angles, paramf, parami):
    c = decode_fbp(paramf, parami)
    tomo = np.asarray(tomo, dtype=np.float32).reshape(c["nslices"], c["nangles"], c["nrays"])
    q = filter_rows(tomo, paramf, parami)
    N, n, L = c["nrays"], c["npad"], c["L"]
    ang = np.asarray(angles, dtype=np.float32)
    th = ang.astype(np.float64)
    dth = angle_steps(ang)
    scale = np.pi / th[-1] if ang[-1] > np.float32(np.pi) else 1.0
    half = n / 2.0
    I = np.arange(N) + L
    x = (I - half)[None, :]
    y = (I - half)[:, None]
    out = np.zeros((c["nslices"], N, N))
    for a in range(c["nangles"]):
        u = x * np.cos(th[a]) - y * np.sin(th[a]) + half
        T = np.trunc(u).astype(np.int64)
        ok = (T > -1) & (T < n)
        Tc = np.clip(T, 0, n - 1)
        out += np.where(ok[None], q[:, a, :][:, Tc], 0.0) * dth[a]
    return out * scale


def radon(phantom, angles, nrays=None, ax=1.0, ay=1.0):
    ph = np.asarray(phantom, dtype=np.float32)
    if ph.ndim == 2:
        ph = ph[None]
    S, Ny, Nx = ph.shape
    nrays = Nx if nrays is None else nrays
    ax, ay = float(np.float32(ax)), float(np.float32(ay))
    dy = 2 * ay / (Ny - 1)
    th = np.asarray(angles, dtype=np.float32).astype(np.float64)
    # (x+1)/dx, (y+1)/dy with t = -ax + i dx = (i - ci) dx, s = -ay + k dy = (k - ck) dy
    ir = (np.arange(nrays) - (Nx - 1) / 2.0)[:, None]
    kr = (np.arange(Ny) - (Ny - 1) / 2.0)[None, :]
    ry = (ay * (Nx - 1)) / (ax * (Ny - 1))
    rx = (ax * (Ny - 1)) / (ay * (Nx - 1))
    out = np.zeros((S, th.size, nrays))
    for a in range(th.size):
        c, sn = np.cos(th[a]), np.sin(th[a])
        ux = ir * c - kr * (sn * ry) + (Nx - 1) / (2.0 * ax)
        uy = ir * (sn * rx) + kr * c + (Ny - 1) / (2.0 * ay)
        X = np.trunc(ux).astype(np.int64)
        Y = np.trunc(uy).astype(np.int64)
        ok = (X >= 0) & (X < Nx) & (Y >= 0) & (Y < Ny)
        vals = ph[:, np.clip(Y, 0, Ny - 1), np.clip(X, 0, Nx - 1)].astype(np.float64)
        out[:, a, :] = np.where(ok[None], vals, 0.0).sum(axis=-1) * dy
    return out


def em_forward(x, angles, nrays):
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    dt = 2.0 / (nrays - 1)
    th = np.asarray(angles, dtype=np.float32).astype(np.float64)
    ic = (nrays - 1) / 2.0
    ir = (np.arange(nrays) - ic)[:, None]
    kr = (np.arange(nrays) - ic)[None, :]
    out = np.zeros((th.size, nrays))
    for a in range(th.size):
        c, sn = np.cos(th[a]), np.sin(th[a])
        X = np.trunc(ir * c - kr * sn + ic).astype(np.int64)
        Y = np.trunc(ir * sn + kr * c + ic).astype(np.int64)
        ok = (X >= 0) & (X < n) & (Y >= 0) & (Y < n)
        out[a] = np.where(ok, x[np.clip(Y, 0, n - 1), np.clip(X, 0, n - 1)], 0.0).sum(axis=-1) * dt
    return out


def em_backproject(q, angles, n, ones=False):
    q = np.asarray(q, dtype=np.float64)
    A, R = q.shape
    th = np.asarray(angles, dtype=np.float32).astype(np.float64)
    dth = angle_steps(angles)
    rr = (R - 1) / (n - 1)
    ir = (np.arange(n) - (n - 1) / 2.0)[None, :]
    jr = (np.arange(n) - (n - 1) / 2.0)[:, None]
    out = np.zeros((n, n))
    for a in range(A):
        # (x_i cos + y_j sin + 1)/dt, x_i = -1 + i dxy, about the image centre
        c, sn = np.cos(th[a]), np.sin(th[a])
        u = ir * (c * rr) + jr * (sn * rr) + (R - 1) / 2.0
        T = np.trunc(u).astype(np.int64)
        ok = (T > -1) & (T < R)
        v = 1.0 if ones else q[a, np.clip(T, 0, R - 1)]
        out += np.where(ok, v, 0.0) * dth[a]
    return out


def kl(y, v):
    y = np.asarray(y, dtype=np.float64).ravel()
    v = np.asarray(v, dtype=np.float64).ravel()
    ok = (v > 0) & (y >= 0)
    yy, vv = y[ok], v[ok]
    with np.errstate(divide="ignore", invalid="ignore"):
        term = vv - yy + np.where(yy > 0, yy * np.log(np.where(yy > 0, yy, 1.0) / vv), 0.0)
    return float(np.sum(term)), int((~ok).sum())


def eem(tomo, angles, iterations, obj=None):
    tomo = np.asarray(tomo, dtype=np.float32)
    if tomo.ndim == 2:
        tomo = tomo[None]
    S, A, R = tomo.shape
    x = np.ones((S, R, R)) if obj is None else np.array(obj, dtype=np.float64)
    e = float(np.float32(1e-4))
    kls = np.zeros((S, iterations))
    for s in range(S):
        back1 = em_backproject(np.zeros((A, R)), angles, R, ones=True)
        y = tomo[s].astype(np.float64)
        for it in range(iterations):
            v = em_forward(x[s], angles, R)
            kls[s, it] = kl(y, v)[0]
            r = y * np.sign(v) / np.sqrt(e * e + v * v)
            r = np.maximum(r, 0.0)
            x[s] = x[s] * em_backproject(r, angles, R) / back1
    return x, kls


def bst(tomo, paramf, parami):
    """getBSTMultiGPU restated in float64 (bst.cu:22-140,251-312,536-736,
    filters.cu:295-330,369-390, padding.cu:245-268) for tomo [S, A, N]:
    pad (edge replicate, L = pad N / 2) -> optional filter C2C with
    apply(min(k, n-k)/n) exp(-2 pi i c k / n), c = int(offset), unnormalised
    inverse, real part -> mirrored half-rows (weights 4, 2 at index 0, upper
    half zero) -> FFT -> sigma_k = 2 / (n max(min(k, n-k), 1/2)) -> polar to
    Cartesian gather (the reference's weights tfrac = iarc - angle) with the
    checkerboard sign at the shifted index -> unnormalised 2-D inverse FFT ->
    Re / (4 n pixel), cropped at L.  Angles are assumed uniform on [0, pi)."""
    c = decode_fbp(paramf, parami)
    tomo = np.asarray(tomo, dtype=np.float32).reshape(c["nslices"], c["nangles"], c["nrays"])
    S, A, N = tomo.shape
    n, L = c["npad"], c["L"]
    u = pad_rows(tomo, n, L)  # [S, A, n] float64
    if c["ftype"] != 0:
        k = np.arange(n)
        w = np.minimum(k, n - k) / n
        H = filter_apply(c["ftype"], w, c["reg"], c["pag"]) * np.exp(-2j * np.pi * int(c["offset"]) * k / n)
        u = np.real(np.fft.ifft(np.fft.fft(u, axis=-1) * H, axis=-1) * n)
    half = n // 2
    idx = np.arange(half)
    wts = np.where(idx == 0, 2.0, 4.0)
    P = np.zeros((S, 2 * A, n), dtype=np.complex128)
    P[:, :A, :half] = u[:, :, half - 1 - idx] * wts
    P[:, A:, :half] = u[:, :, half + idx] * wts
    P = np.fft.fft(P, axis=-1)
    kk = np.arange(n)
    P *= 2.0 / (n * np.maximum(np.minimum(kk, n - kk), 0.5))
    s = n
    A2 = 2 * A
    tx = np.arange(s)[None, :]
    ty = np.arange(s)[:, None]
    posx, posy = tx - s // 2, ty - s // 2
    rho = n * np.hypot(posx, posy) / s
    angle = A2 * (0.5 * np.arctan2(posy, posx) / np.pi + 0.5)
    irho = np.floor(rho).astype(np.int64)
    iarc = np.floor(angle).astype(np.int64)
    ok = irho < n // 2 - 1
    pfrac = rho - irho
    tfrac = iarc - angle
    ia = iarc % A2
    ua = (ia + 1) % A2
    ir = np.where(ok, irho, 0)
    out = np.zeros((S, N, N))
    sign = np.where((tx % 2) == (ty % 2), 1.0, -1.0)
    for sl in range(S):
        Q = P[sl]
        i0 = Q[ia, ir] * (1 - pfrac) + Q[ia, ir + 1] * pfrac
        i1 = Q[ua, ir] * (1 - pfrac) + Q[ua, ir + 1] * pfrac
        val = np.where(ok, i0 * tfrac + i1 * (1 - tfrac), 0.0) * sign
        C = np.zeros((s, s), dtype=np.complex128)
        C[(ty + s // 2) % s, (tx + s // 2) % s] = val
        img = np.real(np.fft.ifft2(C)) * (s * s) / (4.0 * n * c["pixel"])
        out[sl] = img[L:L + N, L:L + N]
    return out
