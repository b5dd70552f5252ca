"""Seeded synthetic SLC stacks for tests and benchmarks.

Model (the coherence model the reference's integration test simulates,
tests/evd/test_evd.py:14-71, with the workload parameters used here):
  * dates t_n = 12 n days;
  * true coherence C_ij = (g0 - ginf) exp(-|t_i - t_j| / tau) + ginf,
    C_ii = 1, with g0 - ginf = 0.6, ginf = 0.2, tau = 24 days;
  * deformation phase phi_n(col) = n * ramp * col / (cols - 1)
    (a linear ramp of 0..ramp rad across the image per date step);
  * samples z = D_phi C^{1/2} w with w ~ CN(0, I) and the eigen square root
    whose eigenvalues below 1e-3 are floored to zero (test_evd.py:14-24);
    D_phi^H w has the law of w, so z = D_phi (C^{1/2} w).
  * optional amplitude patches: amplitudes replaced by sqrt(Gamma(k_p,
    theta_p)) draws, k_p ~ U[1, 4] and theta_p log-uniform over one decade,
    drawn per square patch; the phase keeps the coherent model.
"""
import numpy as np


def dates(n, spacing=12.0):
    return spacing * np.arange(n, dtype=np.float64)


def coherence_matrix(n, spacing=12.0, g0=0.8, ginf=0.2, tau=24.0):
    t = dates(n, spacing)
    dt = np.abs(t[:, None] - t[None, :])
    C = (g0 - ginf) * np.exp(-dt / tau) + ginf
    np.fill_diagonal(C, 1.0)
    return C


def coherence_sqrt(C):
    w, v = np.linalg.eigh(C)
    w = np.where(w < 1e-3, 0.0, w)
    return (v * np.sqrt(w)) @ v.conj().T


def ramp_phase(n, cols, ramp=3.0):
    x = np.arange(cols, dtype=np.float64) / max(cols - 1, 1)
    return np.arange(n, dtype=np.float64)[:, None] * ramp * x[None, :]   # [n, cols]


def stack(n, rows, cols, seed, ramp=3.0, patches=0, coh=None):
    """complex64 [n, rows, cols] on the host (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C = coherence_matrix(n) if coh is None else coh
    L = coherence_sqrt(C).astype(np.complex128)
    w = (rng.standard_normal((n, rows * cols)) + 1j * rng.standard_normal((n, rows * cols))) / np.sqrt(2.0)
    u = (L @ w).reshape(n, rows, cols)
    ph = ramp_phase(n, cols, ramp)
    z = u * np.exp(1j * ph)[:, None, :]
    if patches:
        amp = patch_amplitudes(rng, n, rows, cols, patches)
        z = amp * np.exp(1j * np.angle(z))
    return z.astype(np.complex64)


def patch_amplitudes(rng, n, rows, cols, patch):
    pr, pc = (rows + patch - 1) // patch, (cols + patch - 1) // patch
    k = rng.uniform(1.0, 4.0, size=(pr, pc))
    theta = 10.0 ** rng.uniform(0.0, 1.0, size=(pr, pc))
    kk = np.repeat(np.repeat(k, patch, 0), patch, 1)[:rows, :cols]
    th = np.repeat(np.repeat(theta, patch, 0), patch, 1)[:rows, :cols]
    g = rng.gamma(kk[None].repeat(n, 0), th[None].repeat(n, 0))
    return np.sqrt(g)


def stack_torch(n, rows, cols, seed, device="cuda", ramp=3.0, patches=0):
    """complex64 [n, rows, cols] generated on the device (torch generator),
    same model as ``stack``; used where the host generator would dominate
    set-up time.  Not bit-identical to ``stack``."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    L = torch.from_numpy(coherence_sqrt(coherence_matrix(n))).to(device=device, dtype=torch.complex64)
    out = torch.empty((n, rows, cols), dtype=torch.complex64, device=device)
    ph = torch.from_numpy(ramp_phase(n, cols, ramp)).to(device=device, dtype=torch.float32)
    step = max(1, (1 << 24) // max(cols, 1))
    for r0 in range(0, rows, step):
        r1 = min(rows, r0 + step)
        m = (r1 - r0) * cols
        w = torch.complex(torch.randn((n, m), generator=g, device=device),
                          torch.randn((n, m), generator=g, device=device)) * (0.5 ** 0.5)
        u = (L @ w).reshape(n, r1 - r0, cols)
        out[:, r0:r1] = u * torch.polar(torch.ones_like(ph), ph)[:, None, :]
    if patches:
        pr, pc = (rows + patches - 1) // patches, (cols + patches - 1) // patches
        k = torch.rand((pr, pc), generator=g, device=device) * 3.0 + 1.0
        th = 10.0 ** torch.rand((pr, pc), generator=g, device=device)
        kk = k.repeat_interleave(patches, 0).repeat_interleave(patches, 1)[:rows, :cols]
        tt = th.repeat_interleave(patches, 0).repeat_interleave(patches, 1)[:rows, :cols]
        for b in range(n):
            gam = torch._standard_gamma(kk.contiguous()) * tt
            out[b] = torch.sqrt(gam) * torch.exp(1j * torch.angle(out[b]))
    return out


def full_mask(rows, cols, Nx, Ny):
    """Bitmask with every window bit set: uint32 [rows, cols, nulong]."""
    W = (2 * Nx + 1) * (2 * Ny + 1)
    nul = (W + 31) // 32
    words = np.zeros(nul, np.uint32)
    for f in range(W):
        words[f >> 5] |= np.uint32(1 << (f & 31))
    return np.broadcast_to(words, (rows, cols, nul)).copy()


def random_mask_torch(rows, cols, Nx, Ny, fill, seed, device="cuda"):
    """SHP bitmask int32 [rows, cols, nulong] on the device: the centre bit
    always set, every other window bit set with probability `fill`."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    W = (2 * Nx + 1) * (2 * Ny + 1)
    nul = (W + 31) // 32
    centre = Ny * (2 * Nx + 1) + Nx
    out = torch.empty((rows, cols, nul), dtype=torch.int32, device=device)
    shifts = torch.arange(32, device=device, dtype=torch.int64)
    step = max(1, (1 << 22) // max(cols * W, 1) + 1)
    for r0 in range(0, rows, step):
        r1 = min(rows, r0 + step)
        bits = torch.rand((r1 - r0, cols, nul * 32), generator=g, device=device) < fill
        bits[..., W:] = False
        bits[..., centre] = True
        words = (bits.view(r1 - r0, cols, nul, 32).to(torch.int64) << shifts).sum(-1)
        out[r0:r1] = (words - ((words >> 31) << 32)).to(torch.int32)   # two's-complement bit pattern
    return out
