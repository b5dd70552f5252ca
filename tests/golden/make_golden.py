#!/usr/bin/env python3
"""Generate the golden vectors in this folder from the reference's own
header-only kernels, compiled by oracle/Makefile into oracle/_ref:
  * KS2sample::test p-values (src/nmap/KS2sample.hpp) on seeded sample pairs,
    including tied values;
  * KS2sample::KolmogorovProb on a grid of z;
  * AD2unique::test p-values (src/nmap/AD2unique.hpp) on seeded pairs,
    including tied values;
  * Ulongmask bit positions (include/fringe/ulongmask.hpp) for the two window
    sizes tests/bitmask/test_bits.cpp exercises (5x5 and 15x15);
  * EVWorker (include/fringe/EigenLapack.hpp) on the 3x3 known-answer matrix
    of tests/eigen/test_eig.cpp and on seeded Hermitian matrices.
Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libfringe_ref.so is missing (needs the reference tree to build)")
    rng = np.random.Generator(np.random.PCG64(2024))
    # KS pairs
    ns, A, B, P = [], [], [], []
    for n in (8, 10, 16, 25, 30, 40, 50, 64, 96):
        for t in range(40):
            a = rng.rayleigh(size=n) * rng.uniform(0.5, 2.0)
            b = rng.rayleigh(size=n) * rng.uniform(0.5, 2.0)
            if t % 4 == 0:   # ties inside and across the samples
                a, b = np.round(a * 3) / 3, np.round(b * 3) / 3
            a = np.sort(a.astype(np.float32))
            b = np.sort(b.astype(np.float32))
            pa = np.zeros(96, np.float32); pb = np.zeros(96, np.float32)
            pa[:n] = a; pb[:n] = b
            ns.append(n); A.append(pa); B.append(pb)
            P.append(R.ref_ks2_test(a, b, n))
    z = np.linspace(0.0, 7.0, 701)
    kp = np.array([R.ref_kolmogorov_prob(float(v)) for v in z])
    np.savez_compressed(os.path.join(OUT, "ks2.npz"), n=np.array(ns, np.int32), a=np.array(A), b=np.array(B),
             p=np.array(P), z=z, kprob=kp)
    # bit layout
    bits = {}
    for (Ny, Nx) in ((2, 2), (7, 7), (5, 7)):
        nul = int(np.ceil((2 * Ny + 1) * (2 * Nx + 1) / 32))
        words = []
        for ii in range(-Ny, Ny + 1):
            for jj in range(-Nx, Nx + 1):
                w = np.zeros(nul, np.uint32)
                R.ref_setbit(w, Ny, Nx, ii, jj)
                words.append(w.copy())
        bits[f"w_{Ny}_{Nx}"] = np.array(words)
    np.savez(os.path.join(OUT, "bitmask.npz"), **bits)
    # eigen known answer + seeded Hermitian matrices
    def eig(Ac, idx):
        a = np.asfortranarray(Ac, np.complex128).ravel(order="F").view(np.float64).copy()
        import ctypes
        val = ctypes.c_double()
        vec = np.zeros(2 * Ac.shape[0])
        info = R.ref_eigen_index(a, Ac.shape[0], idx, 1, ctypes.byref(val), vec)
        return info, val.value, vec.view(np.complex128)
    kat = np.array([[3, -2, 4], [-2, 8, 2], [4, 2, 3]], np.complex128)
    out = {"kat": kat}
    for idx in (1, 2, 3):
        info, val, vec = eig(kat, idx)
        out[f"kat_val_{idx}"] = val
        out[f"kat_vec_{idx}"] = vec
    inv = np.asfortranarray(kat + 2 * np.eye(3), np.complex128).ravel(order="F").view(np.float64).copy()
    out["kat_inv_info"] = R.ref_pd_inverse(inv, 3)
    out["kat_inv"] = inv.view(np.complex128).reshape((3, 3), order="F")
    for n in (10, 30):
        X = rng.standard_normal((n, 2 * n)) + 1j * rng.standard_normal((n, 2 * n))
        H = X @ X.conj().T / (2 * n)
        out[f"herm_{n}"] = H
        for idx in (1, n):
            info, val, vec = eig(H, idx)
            out[f"herm_{n}_val_{idx}"] = val
            out[f"herm_{n}_vec_{idx}"] = vec
    np.savez(os.path.join(OUT, "eigen.npz"), **out)
    # AD2unique::test p-values on seeded pairs (own generator, so the files
    # above do not change when this part does)
    rng2 = np.random.Generator(np.random.PCG64(2025))
    ns, A, B, P = [], [], [], []
    for n in (8, 10, 16, 30, 48, 64, 96):
        for t in range(40):
            a = rng2.gamma(2.0, size=n) * rng2.uniform(0.5, 2.0)
            b = rng2.gamma(2.0, size=n) * rng2.uniform(0.5, 2.0)
            if t % 4 == 0:   # ties inside and across the samples
                a, b = np.round(a * 3) / 3, np.round(b * 3) / 3
            a = np.sort(a.astype(np.float32))
            b = np.sort(b.astype(np.float32))
            pa = np.zeros(96, np.float32); pb = np.zeros(96, np.float32)
            pa[:n] = a; pb[:n] = b
            ns.append(n); A.append(pa); B.append(pb)
            P.append(R.ref_ad2_test(a, b, n))
    np.savez_compressed(os.path.join(OUT, "ad2.npz"), n=np.array(ns, np.int32), a=np.array(A), b=np.array(B),
                        p=np.array(P))
    print("golden vectors written to", OUT)


if __name__ == "__main__":
    main()
