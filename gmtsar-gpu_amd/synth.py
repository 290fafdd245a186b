"""Synthetic SAR SLC pairs with known offsets (the reference ships no data).

Speckle model: m = (g_r + i g_i)/sqrt(2) * exp(sigma*n - sigma^2/2), sigma = 0.5,
g, n iid N(0,1), scaled so I and Q have std ~ 500, rounded and clipped to int16.
The aligned image is gamma * shift(m) + sqrt(1 - gamma^2) * w with w complex
Gaussian at matched power.  Sign convention: aligned(x) = master(x - D), so a
feature at master column c appears at aligned column c + D and xcorr reports
dx = D (likewise dy).

Images are returned as (lines, samples, 2) int16 arrays (I, Q interleaved), the
layout of a GMTSAR SLC file (read_xcorr_data.c:11-29).
"""
import numpy as np

SIGMA_LN = 0.5
STD_IQ = 500.0


def speckle(shape, rng):
    g = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)
    tex = np.exp(SIGMA_LN * rng.standard_normal(shape) - SIGMA_LN ** 2 / 2.0)
    return g * tex


def fourier_shift(img, dx, dy):
    """Shift a complex image so that out(x) = img(x - dx) (periodic, band-limited)."""
    ny, nx = img.shape
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    F = np.fft.fft2(img)
    F *= np.exp(-2j * np.pi * (kx * dx + ky * dy))
    return np.fft.ifft2(F)


def to_cint16(z, scale):
    out = np.empty(z.shape + (2,), np.int16)
    out[..., 0] = np.clip(np.rint(z.real * scale), -32768, 32767)
    out[..., 1] = np.clip(np.rint(z.imag * scale), -32768, 32767)
    return out


def constant_shift_pair(n_range, n_azimuth, dx, dy, coherence, seed=1, noise_seed=2):
    """Master/aligned pair with a constant (dx, dy) shift (the small parity
    workload uses 1024 x 2048, dx = 3.37, dy = -12.81, coherence 0.8)."""
    rng = np.random.default_rng(seed)
    m = speckle((n_azimuth, n_range), rng)
    s = fourier_shift(m, dx, dy)
    rngw = np.random.default_rng(noise_seed)
    pw = np.sqrt(np.mean(np.abs(m) ** 2))
    w = (rngw.standard_normal(m.shape) + 1j * rngw.standard_normal(m.shape)) / np.sqrt(2.0) * pw
    s = coherence * s + np.sqrt(1.0 - coherence ** 2) * w
    scale = STD_IQ / np.sqrt(np.mean(np.abs(m) ** 2) / 2.0)
    return to_cint16(m, scale), to_cint16(s, scale)


def write_slc(path, img):
    np.ascontiguousarray(img, dtype="<i2").tofile(path)


def write_prm(path, slc_file, num_rng_bins, num_lines, rshift=0, ashift=0, prf=1679.9):
    """Minimal PRM with the keys xcorr reads (sio_struct.c:141-222)."""
    with open(path, "w") as f:
        f.write("SLC_file = %s\n" % slc_file)
        f.write("num_rng_bins = %d\n" % num_rng_bins)
        f.write("num_patches = 1\n")
        f.write("num_valid_az = %d\n" % num_lines)
        f.write("rshift = %d\n" % rshift)
        f.write("ashift = %d\n" % ashift)
        f.write("PRF = %.6f\n" % prf)


def strip_pair_torch(n_range, n_azimuth, dx_fn, dy_fn, coh_range=(0.3, 0.95), block=512, margin=64,
                     seed=7, device="cuda"):
    """Full-strip pair generated on the GPU (the strip workload: 10800 x 27648,
    dx = 2.1 + 1e-4 x, dy = -7.6 + 3e-4 y, coherence per 512 x 512 block uniform
    in [0.3, 0.95]).  Each block is Fourier-shifted by the offset at its centre
    (per-block Fourier resampling with a `margin`-sample apron).  Returns
    (master, slave) as (lines, samples, 2) int16 tensors on `device`."""
    import math

    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    ny, nx = n_azimuth, n_range

    def randn(*shape):
        return torch.randn(*shape, generator=g, device=device, dtype=torch.float32)

    m = torch.complex(randn(ny, nx), randn(ny, nx)) / math.sqrt(2.0)
    m = m * torch.exp(SIGMA_LN * randn(ny, nx) - SIGMA_LN ** 2 / 2.0)
    power = float((m.real.square() + m.imag.square()).mean())
    T = block + 2 * margin
    nby, nbx = -(-ny // block), -(-nx // block)
    mp = torch.zeros(nby * block + 2 * margin, nbx * block + 2 * margin, dtype=torch.complex64, device=device)
    mp[margin:margin + ny, margin:margin + nx] = m
    s = torch.empty(nby * block, nbx * block, dtype=torch.complex64, device=device)
    kx = torch.fft.fftfreq(T, device=device)[None, None, :]
    ky = torch.fft.fftfreq(T, device=device)[None, :, None]
    coh = coh_range[0] + (coh_range[1] - coh_range[0]) * torch.rand(nby, nbx, generator=g, device=device)
    xc = torch.arange(nbx, device=device, dtype=torch.float64) * block + block / 2.0
    for by in range(nby):
        strip = mp[by * block: by * block + T]
        tiles = strip.unfold(1, T, block).permute(1, 0, 2)            # (nbx, T, T)
        yc = by * block + block / 2.0
        dx = torch.as_tensor([dx_fn(float(v)) for v in xc], device=device, dtype=torch.float32)[:, None, None]
        dy = float(dy_fn(yc))
        F = torch.fft.fft2(tiles)
        F = F * torch.exp(-2j * math.pi * (kx * dx + ky * dy)).to(torch.complex64)
        sh = torch.fft.ifft2(F)[:, margin:margin + block, margin:margin + block]   # (nbx, block, block)
        gam = coh[by][:, None, None]
        w = torch.complex(randn(nbx, block, block), randn(nbx, block, block)) * math.sqrt(power / 2.0)
        sh = gam * sh + torch.sqrt(1.0 - gam * gam) * w
        s[by * block:(by + 1) * block] = sh.permute(1, 0, 2).reshape(block, nbx * block)
    s = s[:ny, :nx]
    scale = STD_IQ / math.sqrt(power / 2.0)

    def q(z):
        out = torch.empty(ny, nx, 2, dtype=torch.int16, device=device)
        out[..., 0] = torch.clamp(torch.round(z.real * scale), -32768, 32767).to(torch.int16)
        out[..., 1] = torch.clamp(torch.round(z.imag * scale), -32768, 32767).to(torch.int16)
        return out

    return q(m), q(s)


def sinusoid_field(rng, n=8, amp=15.0, wl=(133.0, 667.0)):
    """Smooth random offset field: sum of n random 2-D sinusoids per component,
    total amplitude <= amp px, wavelengths in wl px (2-10 km at 15 m pixels).
    The returned f(x, y) -> (dx, dy) takes numpy arrays or torch tensors."""
    comps = []
    for _ in range(2):
        a = rng.uniform(0.2, 1.0, n)
        a *= amp / a.sum()
        lam = rng.uniform(wl[0], wl[1], n)
        th = rng.uniform(0, 2 * np.pi, n)
        ph = rng.uniform(0, 2 * np.pi, n)
        comps.append((a, 2 * np.pi / lam * np.cos(th), 2 * np.pi / lam * np.sin(th), ph))

    def f(x, y):
        sin = np.sin if isinstance(x, np.ndarray) or np.isscalar(x) else type(x).sin
        out = []
        for a, kx, ky, ph in comps:
            v = 0.0 * x
            for i in range(n):
                v = v + float(a[i]) * sin(float(kx[i]) * x + float(ky[i]) * y + float(ph[i]))
            out.append(v)
        return out[0], out[1]
    return f


SINC_TAPS = 32          # windowed-sinc resampling kernel: taps -15 .. 16 around floor(position)
KAISER_BETA = 6.0       # 3.9 % error power on full-band speckle against an exact Fourier shift


def _resample_1d(torch, src, pos, axis):
    """Samples of the complex image src at fractional positions pos along
    `axis` (0: lines, 1: samples), the other coordinate being the output's
    own; Kaiser-windowed sinc of SINC_TAPS taps, zero outside the image."""
    import math
    n = src.shape[axis]
    i0 = torch.floor(pos)
    fr = pos - i0
    i0 = i0.to(torch.int64)
    out = torch.zeros(pos.shape, dtype=src.dtype, device=src.device)
    half = SINC_TAPS // 2
    den = float(np.i0(KAISER_BETA))
    for k in range(-half + 1, half + 1):
        t = fr - k                                          # distance from tap to position
        w = torch.sinc(t) * torch.special.i0(KAISER_BETA * torch.sqrt(torch.clamp(1 - (t / half) ** 2, min=0))) / den
        idx = i0 + k
        ok = (idx >= 0) & (idx < n)
        idx = idx.clamp(0, n - 1)
        if axis == 0:
            v = torch.gather(src, 0, idx)
        else:
            v = torch.gather(src, 1, idx)
        out += torch.where(ok, w, torch.zeros_like(w)).to(src.dtype) * v
    return out


def field_pair_torch(n_range, n_azimuth, field, coherence, seed=11, device="cuda", rows=256):
    """Pair with a smooth, spatially varying offset field, resampled per sample
    on the GPU.  The slave is the master resampled in two separable passes
    (windowed sinc, SINC_TAPS taps): first along azimuth, s1(x, y) =
    m(x, y - dy(x, y)), then along range, s(x, y) = s1(x - dx(x, y), y), and
    mixed with noise at coherence(x, y).  The displacement of output sample
    (x, y) is therefore exactly (dx(x, y), dy(x - dx(x, y), y))
    (field_displacement).  field and coherence take torch tensors.  Returns
    (master, slave) as (lines, samples, 2) int16 tensors on `device`."""
    import math

    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    ny, nx = n_azimuth, n_range

    def randn(*shape):
        return torch.randn(*shape, generator=g, device=device, dtype=torch.float32)

    m = torch.complex(randn(ny, nx), randn(ny, nx)) / math.sqrt(2.0)
    m = m * torch.exp(SIGMA_LN * randn(ny, nx) - SIGMA_LN ** 2 / 2.0)
    power = float((m.real.square() + m.imag.square()).mean())
    xs = torch.arange(nx, device=device, dtype=torch.float32)
    s1 = torch.empty_like(m)
    for y0 in range(0, ny, rows):
        yy = torch.arange(y0, min(ny, y0 + rows), device=device, dtype=torch.float32)
        Y, X = torch.meshgrid(yy, xs, indexing="ij")
        _, dy = field(X, Y)
        lo, hi = max(0, y0 - 64), min(ny, y0 + rows + 64)      # |dy| <= 15 px plus the taps
        s1[y0:y0 + len(yy)] = _resample_1d(torch, m[lo:hi], Y - dy - lo, 0)
    s = torch.empty_like(m)
    for y0 in range(0, ny, rows):
        yy = torch.arange(y0, min(ny, y0 + rows), device=device, dtype=torch.float32)
        Y, X = torch.meshgrid(yy, xs, indexing="ij")
        dx, _ = field(X, Y)
        gam = torch.clamp(coherence(X, Y), 0.0, 1.0)
        v = _resample_1d(torch, s1[y0:y0 + len(yy)], X - dx, 1)
        w = torch.complex(randn(*v.shape), randn(*v.shape)) * math.sqrt(power / 2.0)
        s[y0:y0 + len(yy)] = gam * v + torch.sqrt(1.0 - gam * gam) * w
    scale = STD_IQ / math.sqrt(power / 2.0)

    def q(z):
        out = torch.empty(ny, nx, 2, dtype=torch.int16, device=device)
        out[..., 0] = torch.clamp(torch.round(z.real * scale), -32768, 32767).to(torch.int16)
        out[..., 1] = torch.clamp(torch.round(z.imag * scale), -32768, 32767).to(torch.int16)
        return out

    return q(m), q(s)


def field_displacement(field, x, y):
    """Displacement of slave sample (x, y) in a field_pair_torch pair."""
    dx, _ = field(x, y)
    _, dy = field(x - dx, y)
    return dx, dy


def window_truth(field, x, y, half, k=8):
    """Offset a window centred on (x, y) with a template of 2 half x 2 half
    samples should report: the mean displacement over the template, on a
    k x k grid of sample centres.  Also returns the spread of the displacement
    over the template (max - min per component)."""
    t = (np.arange(k) + 0.5) / k * (2 * half) - half
    gx = np.asarray(x, np.float64)[:, None, None] + t[None, None, :]
    gy = np.asarray(y, np.float64)[:, None, None] + t[None, :, None]
    gx, gy = np.broadcast_arrays(gx, gy)
    dx, dy = field_displacement(field, gx, gy)
    return (dx.mean(axis=(1, 2)), dy.mean(axis=(1, 2)),
            dx.max(axis=(1, 2)) - dx.min(axis=(1, 2)), dy.max(axis=(1, 2)) - dy.min(axis=(1, 2)))


def smooth_coherence(rng, lo=0.15, hi=0.95, n=6, wl=(1000.0, 6000.0)):
    """Spatially correlated coherence in [lo, hi]: a sum of random sinusoids
    mapped through a logistic."""
    lam = rng.uniform(wl[0], wl[1], n)
    th = rng.uniform(0, 2 * np.pi, n)
    ph = rng.uniform(0, 2 * np.pi, n)

    def f(x, y):
        mod = np if isinstance(x, np.ndarray) or np.isscalar(x) else type(x)
        v = 0.0 * x
        for i in range(n):
            v = v + mod.sin(float(2 * np.pi / lam[i]) * (float(np.cos(th[i])) * x + float(np.sin(th[i])) * y)
                            + float(ph[i]))
        return lo + (hi - lo) / (1.0 + mod.exp(-1.5 * v))
    return f
