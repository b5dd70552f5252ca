"""CPU tests that pin the oracle (test infrastructure) before it is trusted:
against the reference's own header kernels (golden vectors in tests/golden,
and oracle/_ref where it is built), against closed forms and against brute
force on tiny inputs."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from fringe_amd import synthetic

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ---------------------------------------------------------------- KS test --
def test_ks_pvalues_match_reference_golden():
    g = _load("ks2.npz")
    for n, a, b, p in zip(g["n"], g["a"], g["b"], g["p"]):
        assert oracle.ks2_pvalue(a[:n], b[:n]) == p


def test_kolmogorov_prob_matches_reference_golden():
    g = _load("ks2.npz")
    got = np.array([oracle.kolmogorov_prob(z) for z in g["z"]])
    assert np.array_equal(got, g["kprob"])


def _ks_brute(a, b):
    """Exact two-sample KS statistic D*n from the empirical CDFs."""
    xs = np.union1d(a, b)
    fa = np.searchsorted(np.sort(a), xs, side="right")
    fb = np.searchsorted(np.sort(b), xs, side="right")
    return int(np.abs(fa - fb).max())


@pytest.mark.parametrize("n", [5, 10, 30, 64])
def test_ks_statistic_is_exact_ecdf_distance(n):
    rng = np.random.default_rng(n)
    for t in range(300):
        a = np.sort(rng.integers(0, 12 if t % 2 else 10000, n).astype(np.float32))
        b = np.sort(rng.integers(0, 12 if t % 2 else 10000, n).astype(np.float32))
        d = oracle.ks2_statistic(a, b)
        assert abs(d * n - _ks_brute(a, b)) < 1e-9


def test_ks_reference_shim_when_built():
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(7)
    for t in range(500):
        n = int(rng.integers(4, 97))
        a = np.sort(rng.gamma(2.0, size=n).astype(np.float32))
        b = np.sort((rng.gamma(2.0, size=n) * rng.uniform(0.8, 1.3)).astype(np.float32))
        assert oracle.ks2_pvalue(a, b) == R.ref_ks2_test(a, b, n)


def test_ks_symmetric():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = np.sort(rng.rayleigh(size=30).astype(np.float32))
        b = np.sort(rng.rayleigh(size=30).astype(np.float32))
        assert oracle.ks2_pvalue(a, b) == oracle.ks2_pvalue(b, a)


# ------------------------------------------------------------- bit layout --
def test_bit_layout_matches_reference_golden():
    g = _load("bitmask.npz")
    for (Ny, Nx) in ((2, 2), (7, 7), (5, 7)):
        words = g[f"w_{Ny}_{Nx}"]
        k = 0
        for ii in range(-Ny, Ny + 1):
            for jj in range(-Nx, Nx + 1):
                flat = (ii + Ny) * (2 * Nx + 1) + jj + Nx
                exp = np.zeros_like(words[k])
                exp[flat >> 5] = np.uint32(1 << (flat & 31))
                assert np.array_equal(words[k], exp)
                k += 1
    assert oracle.mask_words(5, 5) == 4 and oracle.mask_words(7, 5) == 6 and oracle.mask_words(10, 5) == 8


def test_unpackbits_reads_full_mask_like_reference_test():
    # tests/evd/test_evd.py:114-226: LSB-first unpackbits of the words gives
    # the window row-major, padding bits zero
    Nx, Ny = 5, 5
    w = synthetic.full_mask(1, 1, Nx, Ny)[0, 0]
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")
    W = (2 * Nx + 1) * (2 * Ny + 1)
    assert bits[:W].all() and not bits[W:].any()


# ------------------------------------------------------------------ eigen --
def test_eigen_known_answer():
    g = _load("eigen.npz")
    A = g["kat"]
    exact = sorted([7.0, (7 + 113 ** 0.5) / 2, (7 - 113 ** 0.5) / 2])
    for idx in (1, 2, 3):
        info, val, vec = oracle.eigen_index(A, idx)
        assert info == 0
        assert val == g[f"kat_val_{idx}"]
        # zheevr with ABSTOL=1e-6 (EigenLapack.hpp:145) returns eigenvalues to ~1e-6
        assert abs(val - exact[idx - 1]) < 1e-5
        assert np.allclose(A @ vec, val * vec, atol=1e-5)
    info, inv = oracle.pd_inverse(A + 2 * np.eye(3))
    assert info == 0 and g["kat_inv_info"] == 0
    assert np.allclose(inv, g["kat_inv"]) and np.allclose((A + 2 * np.eye(3)) @ inv, np.eye(3), atol=1e-12)


def test_eigen_random_hermitian_matches_reference_golden():
    g = _load("eigen.npz")
    for n in (10, 30):
        H = g[f"herm_{n}"]
        w = np.linalg.eigvalsh(H)
        for idx in (1, n):
            info, val, vec = oracle.eigen_index(H, idx)
            assert info == 0
            assert abs(val - g[f"herm_{n}_val_{idx}"]) < 1e-12
            assert abs(val - w[idx - 1]) < 2e-6
            ref = g[f"herm_{n}_vec_{idx}"]
            # same eigenvector up to a unit phase
            assert abs(abs(np.vdot(ref, vec)) - 1) < 1e-8


# --------------------------------------------------------- nmap (oracle) --
def _amp(slc):
    """|z| as glibc's hypotf (and std::abs of complex<float>) rounds it: the
    double square root of the exact double sum of squares, rounded to float
    (numpy's complex64 abs can differ in the last bit)."""
    z = slc.astype(np.complex64)
    re, im = z.real.astype(np.float64), z.imag.astype(np.float64)
    return np.sqrt(re * re + im * im).astype(np.float32)


def _nmap_brute(slc, Nx, Ny, prob):
    nb, rows, cols = slc.shape
    amp = _amp(slc)
    valid = np.all((amp != 0) & ~np.isnan(amp), axis=0)
    srt = np.sort(amp, axis=0)
    nul = oracle.mask_words(Nx, Ny)
    wts = np.zeros((rows, cols, nul), np.uint32)
    cnt = np.zeros((rows, cols), np.int32)
    for r in range(rows):
        for c in range(cols):
            if not valid[r, c]:
                continue
            for ii in range(-Ny, Ny + 1):
                for jj in range(-Nx, Nx + 1):
                    rr, cc = r + ii, c + jj
                    if not (0 <= rr < rows and 0 <= cc < cols) or not valid[rr, cc]:
                        continue
                    k = _ks_brute(srt[:, r, c], srt[:, rr, cc])
                    z = k / nb * np.sqrt(nb * nb / (2.0 * nb))
                    if (rr, cc) == (r, c) or oracle.kolmogorov_prob(z) >= prob:
                        f = (ii + Ny) * (2 * Nx + 1) + jj + Nx
                        wts[r, c, f >> 5] |= np.uint32(1 << (f & 31))
                        cnt[r, c] += 1
    return wts, cnt


def test_nmap_oracle_matches_brute_force():
    slc = synthetic.stack(12, 9, 11, seed=11, patches=4)
    slc[3, 4, 5] = 0   # an invalid pixel
    w1, c1 = oracle.nmap(slc, 3, 2, 0.05)
    w2, c2 = _nmap_brute(slc, 3, 2, 0.05)
    assert np.array_equal(w1, w2) and np.array_equal(c1, c2)
    assert c1[4, 5] == 0 and not w1[4, 5].any()


def test_nmap_oracle_matches_brute_force_at_59x19():
    # the reference sequential default window (sequential.py -x 29 -y 9):
    # 1121 positions in 36 words, wider and taller than the image
    slc = synthetic.stack(10, 10, 12, seed=13, patches=4)
    w1, c1 = oracle.nmap(slc, 29, 9, 0.05)
    w2, c2 = _nmap_brute(slc, 29, 9, 0.05)
    assert w1.shape[-1] == 36
    assert np.array_equal(w1, w2) and np.array_equal(c1, c2)


def test_nmap_amplitude_constants_and_mask():
    slc = synthetic.stack(8, 8, 8, seed=12, patches=4)
    mask = np.ones((8, 8), np.uint8)
    mask[2, 2] = 0
    w, c = oracle.nmap(slc, 2, 2, 0.05, mask=mask, amp_const=np.array([2.0, -1, 0, 4, 2, 2, 2, 2]))
    assert c[2, 2] == 0
    # scaling a band by its own constant leaves the map unchanged
    scaled = slc.copy()
    scaled[3] *= 2.0
    w2, c2 = oracle.nmap(scaled, 2, 2, 0.05, amp_const=np.array([1.0, 1, 1, 2, 1, 1, 1, 1]))
    w3, c3 = oracle.nmap(slc, 2, 2, 0.05)
    assert np.array_equal(w2, w3) and np.array_equal(c2, c3)


def test_ad2_pvalues_match_reference_golden():
    g = _load("ad2.npz")
    for n, a, b, p in zip(g["n"], g["a"], g["b"], g["p"]):
        assert oracle.ad2_pvalue(a[:n], b[:n]) == p


def test_ad2_statistic_matches_scipy_anderson_ksamp():
    # without ties the standardised statistic is scipy's k-sample A2 (k = 2,
    # midrank=False), computed there from the textbook formula
    import warnings
    from scipy.stats import anderson_ksamp
    rng = np.random.default_rng(3)
    for n in (8, 10, 30, 64):
        for t in range(40):
            a = np.sort(rng.gamma(2.0, size=n)).astype(np.float32)
            b = np.sort(rng.gamma(2.0, size=n) * rng.uniform(0.6, 1.6)).astype(np.float32)
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                s = anderson_ksamp([a.astype(np.float64), b.astype(np.float64)], midrank=False).statistic
            assert abs(oracle.ad2_statistic(a, b) - s) < 1e-9


def test_ad2_pvalue_at_table_points():
    # the interpolated log-odds pass through every tabulated quantile, where
    # the p-value is the table's upper-tail probability
    t = [-1.1954, -0.2966, 1.9633, 3.7825, 11.8537]
    p = [1 - .00001, 1 - .5, 1 - .95, 1 - .99, 1 - .99999]
    for tt, pp in zip(t, p):
        assert abs(oracle.ad2_pvalue_of(tt) - pp) < 1e-12
    xs = np.linspace(-3, 15, 2001)
    ps = np.array([oracle.ad2_pvalue_of(x) for x in xs])
    assert np.all(np.diff(ps) <= 0)       # decreasing in the statistic


def test_ad2_nmap_follows_the_reference_half_plane_sweep():
    # the reference tests (p, q) with p first for q = p + (ii, jj) in its
    # half plane and sets both bits (nmap.cpp:402-473); ties make the order
    # matter, so quantised amplitudes are used
    R = oracle.ref()
    pv = (lambda a, b: R.ref_ad2_test(a, b, a.size)) if R is not None else oracle.ad2_pvalue
    slc = synthetic.stack(10, 8, 9, seed=13, patches=3)
    slc = (np.round(np.abs(slc) * 3) / 3 * np.exp(1j * np.angle(slc))).astype(np.complex64)
    slc[2, 3, 3] = 0
    Nx, Ny, prob = 2, 2, 0.05
    nb, rows, cols = slc.shape
    amp = _amp(slc)
    valid = np.all(amp != 0, axis=0)
    srt = np.sort(amp, axis=0)
    nul = oracle.mask_words(Nx, Ny)
    w = np.zeros((rows, cols, nul), np.uint32)
    cnt = np.zeros((rows, cols), np.int32)

    def setb(r, c, ii, jj):
        f = (ii + Ny) * (2 * Nx + 1) + jj + Nx
        w[r, c, f >> 5] |= np.uint32(1 << (f & 31))
        cnt[r, c] += 1

    for ii in range(0, Ny + 1):
        for jj in range(0 if ii == 0 else -Nx, Nx + 1):
            for r in range(rows - ii):
                for c in range(max(-jj, 0), min(cols - jj, cols)):
                    rr, cc = r + ii, c + jj
                    if not (valid[r, c] and valid[rr, cc]):
                        continue
                    if (ii, jj) == (0, 0):
                        setb(r, c, 0, 0)
                    elif pv(np.ascontiguousarray(srt[:, r, c]), np.ascontiguousarray(srt[:, rr, cc])) >= prob:
                        setb(r, c, ii, jj)
                        setb(rr, cc, -ii, -jj)
    w1, c1 = oracle.nmap(slc, Nx, Ny, prob, method="AD2")
    assert np.array_equal(w1, w) and np.array_equal(c1, cnt)
    w2, _ = oracle.nmap(slc, Nx, Ny, prob, method="KS2")
    assert not np.array_equal(w1, w2)     # the two tests decide differently here


# ------------------------------------------------------ linking (oracle) --
def _window_samples(slc, wts, r, c, Nx, Ny):
    nb, rows, cols = slc.shape
    Z = []
    for ii in range(-Ny, Ny + 1):
        for jj in range(-Nx, Nx + 1):
            f = (ii + Ny) * (2 * Nx + 1) + jj + Nx
            rr, cc = r + ii, c + jj
            if 0 <= rr < rows and 0 <= cc < cols and (wts[r, c, f >> 5] >> (f & 31)) & 1:
                Z.append(slc[:, rr, cc])
    return np.array(Z, np.complex128)


def test_link_evd_matches_numpy_eigh():
    n, rows, cols, Nx, Ny = 8, 7, 9, 2, 2
    slc = synthetic.stack(n, rows, cols, seed=21)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    out = oracle.link(slc, wts, Nx, Ny, module="evd", method="EVD", want_gap=True)
    for r in range(rows):
        for c in range(cols):
            Z = _window_samples(slc, wts, r, c, Nx, Ny)
            G = Z.T @ Z.conj()
            d = np.sqrt(np.real(np.diag(G)))
            gam = G / np.outer(d, d)
            w, V = np.linalg.eigh(gam)
            v = V[:, -1]
            ph = np.exp(1j * (np.angle(v) - np.angle(v[0])))
            assert np.abs(np.angle(out["phase"][:, r, c] * ph.conj())).max() < 1e-6
            # diagnostics: gap over spread and the smallest unit-eigenvector
            # component relative to the RMS one (zheevr runs with the
            # reference's ABSTOL = 1e-6)
            assert abs(out["gap"][r, c] - (w[-1] - w[-2]) / (w[-1] - w[0])) < 1e-6
            assert abs(out["vmin"][r, c] - np.sqrt(n) * np.abs(v).min()) < 1e-6
            # temporal coherence from its definition (evd.cpp:770-786)
            iu = np.triu_indices(n, 1)
            e = np.exp(1j * (np.angle(gam[iu]) - np.angle(ph[iu[0]]) + np.angle(ph[iu[1]])))
            assert abs(np.abs(e.sum()) / len(e) - out["tcorr"][r, c]) < 1e-6
            comp = np.sum(slc[:, r, c] * ph.conj()) / n
            assert abs(comp - out["comp"][r, c]) < 1e-5 * max(1.0, abs(comp))


def _rank1_stack(n, rows, cols, theta, C, seed):
    """Samples whose sample coherence converges to D_theta C D_theta^H."""
    rng = np.random.default_rng(seed)
    L = np.linalg.cholesky(C)
    K = rows * cols
    w = (rng.standard_normal((n, K)) + 1j * rng.standard_normal((n, K))) / np.sqrt(2)
    u = (L @ w) * np.exp(1j * theta)[:, None]
    return u.reshape(n, rows, cols).astype(np.complex64)


def test_closed_form_emi_and_evd_recover_phase():
    # For Gamma = D_theta C D_theta^H with C real, positive, unit diagonal:
    # EMI's |Gamma|^-1 o Gamma has eigenvalue 1 with eigenvector exp(i theta)
    # and EVD's top eigenvector is D_theta u with u > 0 (Perron-Frobenius),
    # so both return theta exactly and tcorr = 1.
    n = 6
    t = 12.0 * np.arange(n)
    C = 0.7 * np.exp(-np.abs(t[:, None] - t[None, :]) / 36.0) + 0.3
    theta = np.array([0.0, 0.4, -1.1, 2.0, 2.9, -2.5])
    # build a single pixel whose window samples give Gamma exactly:
    # take K = n samples z_k = D_theta L e_k so sum z z^H = D C D^H
    L = np.linalg.cholesky(C)
    Zs = (np.exp(1j * theta)[:, None] * L).T   # rows = samples
    rows, cols, Nx, Ny = 1, n, n, 0
    slc = np.zeros((n, rows, cols), np.complex64)
    for k in range(n):
        slc[:, 0, k] = Zs[k]
    wts = np.zeros((rows, cols, oracle.mask_words(Nx, Ny)), np.uint32)
    # centre pixel k0 sees every sample of its row
    for c in range(cols):
        for jj in range(-Nx, Nx + 1):
            if 0 <= c + jj < cols:
                f = jj + Nx
                wts[0, c, f >> 5] |= np.uint32(1 << (f & 31))
    k0 = 0
    for module, method in (("phase_link", "MLE"), ("evd", "EVD"), ("evd", "MLE")):
        out = oracle.link(slc, wts, Nx, Ny, module=module, method=method)
        exp = np.exp(1j * (theta - theta[0]))
        got = out["phase"][:, 0, k0]
        assert np.abs(np.angle(got * exp.conj())).max() < 2e-6, (module, method)
        assert abs(out["tcorr"][0, k0] - 1.0) < 1e-6


def test_codes_untouched_and_failed_pixels():
    n, rows, cols, Nx, Ny = 10, 6, 6, 1, 1
    slc = synthetic.stack(n, rows, cols, seed=5)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    wts[2, 2, :] = 0                                    # centre bit cleared: untouched
    out = oracle.link(slc, wts, Nx, Ny, module="evd", method="MLE")
    assert out["tcorr"][2, 2] == 0 and not out["phase"][:, 2, 2].any()
    # 9 samples < 10 dates: Gamma is rank deficient -> evd MLE code -2
    assert np.all(out["tcorr"][1:-1, 1:-1][out["tcorr"][1:-1, 1:-1] != 0] == -2)
    # phase_link honours minNeighbors (phase_link.cpp:524), evd does not (evd.cpp:566)
    o2 = oracle.link(slc, wts, Nx, Ny, module="phase_link", method="MLE", min_neighbors=10)
    assert np.all(o2["tcorr"] == 0)
    o3 = oracle.link(slc, wts, Nx, Ny, module="evd", method="EVD")
    inner = o3["tcorr"][1:-1, 1:-1].copy()
    inner[1, 1] = 1.0                                    # pixel (2, 2) stays untouched
    assert np.all(inner != 0) and o3["tcorr"][2, 2] == 0


def test_stbas_bandwidth_error():
    slc = synthetic.stack(6, 4, 4, seed=1)
    wts = synthetic.full_mask(4, 4, 1, 1)
    assert oracle.link(slc, wts, 1, 1, module="evd", method="STBAS", bandwidth=0)["status"] == 101
    assert oracle.link(slc, wts, 1, 1, module="evd", method="STBAS", bandwidth=5)["status"] == 101
    assert oracle.link(slc, wts, 1, 1, module="evd", method="STBAS", bandwidth=2)["status"] == 0


def test_reference_date_and_per_date_scaling():
    n, rows, cols, Nx, Ny = 8, 6, 7, 2, 1
    slc = synthetic.stack(n, rows, cols, seed=8)
    wts = synthetic.full_mask(rows, cols, Nx, Ny)
    a = oracle.link(slc, wts, Nx, Ny, module="phase_link", mini_stack_count=3)
    assert np.all(a["phase"][2][a["tcorr"] > 0] == 1 + 0j)
    scaled = slc * np.linspace(0.5, 3.0, n).astype(np.float32)[:, None, None]
    b = oracle.link(scaled, wts, Nx, Ny, module="phase_link", mini_stack_count=3)
    ok = a["tcorr"] > 0
    assert np.abs(np.angle(a["phase"] * b["phase"].conj()))[:, ok].max() < 1e-5
    assert np.abs(a["tcorr"] - b["tcorr"])[ok].max() < 1e-6


# ------------------------------------------------------------- despeck --
def test_despeck_oracle_closed_forms():
    rows, cols, Nx, Ny = 9, 11, 2, 1
    full = synthetic.full_mask(rows, cols, Nx, Ny)
    # a constant amplitude field stays constant (exact: 2.0 sums exactly)
    z = np.full((rows, cols), 2.0 * np.exp(0.3j), np.complex64)
    amp = oracle.despeck(z, full, Nx, Ny)
    assert np.allclose(amp.real, np.abs(z).astype(np.float32), rtol=0, atol=2e-7) and not amp.imag.any()
    # with only the centre bit, every output is the pixel's own value
    centre = np.zeros_like(full)
    f = Ny * (2 * Nx + 1) + Nx
    centre[..., f >> 5] = np.uint32(1 << (f & 31))
    s = synthetic.stack(2, rows, cols, seed=7)
    a, b = s[0].astype(np.complex64), s[1].astype(np.complex64)
    assert np.array_equal(oracle.despeck(a, centre, Nx, Ny).real, _amp(a))
    assert np.allclose(oracle.despeck(a, centre, Nx, Ny, b2=b), a * np.conj(b), rtol=1e-6)
    coh = oracle.despeck(a, centre, Nx, Ny, b2=b, coherence=True)
    assert np.allclose(np.abs(coh), 1.0, atol=1e-6)
    # the coherence of a band with itself is one
    assert np.allclose(oracle.despeck(a, full, Nx, Ny, b2=a, coherence=True), 1.0, atol=1e-6)
    # a sparse mask averages exactly the selected, in-image neighbours
    rng = np.random.default_rng(1)
    sparse = full & rng.integers(0, 2 ** 32, full.shape, dtype=np.uint64).astype(np.uint32)
    sparse[..., f >> 5] |= np.uint32(1 << (f & 31))
    got = oracle.despeck(a, b2=b, wts=sparse, Nx=Nx, Ny=Ny)
    ifg = (a.astype(np.complex128) * np.conj(b.astype(np.complex128)))
    for r in range(rows):
        for c in range(cols):
            acc, k = 0, 0
            for ii in range(-Ny, Ny + 1):
                for jj in range(-Nx, Nx + 1):
                    ff = (ii + Ny) * (2 * Nx + 1) + jj + Nx
                    if 0 <= r + ii < rows and 0 <= c + jj < cols and (sparse[r, c, ff >> 5] >> (ff & 31)) & 1:
                        acc += ifg[r + ii, c + jj]
                        k += 1
            assert abs(got[r, c] - acc / k) <= 1e-5 * max(1.0, abs(acc / k))


# -------------------------------------------------------------- calamp --
def test_calamp_oracle_closed_forms():
    rows, cols = 6, 7
    z = np.zeros((3, rows, cols), np.complex64)
    z[0] = 3.0 * np.exp(1j * np.linspace(0, 2, rows * cols)).reshape(rows, cols)
    z[1, :2] = 2.0j
    z[1, 2, 3] = np.nan
    z[2] = 5.0
    mask = np.ones((rows, cols), np.uint8)
    mask[:, 0] = 0
    c = oracle.calamp(z, mask)
    assert abs(c[0] - 3.0) < 1e-6           # |3 e^{i t}| rounded to float, averaged
    assert c[1] == 2.0                       # zeros and NaN are not valid pixels
    assert c[2] == 5.0
    assert np.array_equal(oracle.calamp(z, np.zeros_like(mask)), np.ones(3))   # nothing valid -> 1.0
    s = synthetic.stack(2, rows, cols, seed=9)
    want = [float(np.mean(_amp(s[b])[mask > 0].astype(np.float64))) for b in range(2)]
    assert np.allclose(oracle.calamp(s, mask), want, rtol=1e-12)
