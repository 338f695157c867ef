"""Shared test helpers: meshes, perturbed states, comparison norms."""
import numpy as np

from adh_kraken_amd.sw2 import FLUME_C1, FlumeSpec, flume_mesh


def small_flume(npx=11, npy=5, nbumps=3, seed=7, q=1.0, tail=1.0, flat=1):
    spec = FlumeSpec(xmin=0.0, xmax=float(npx - 1), ymin=0.0, ymax=float(npy - 1), npx=npx, npy=npy,
                     slope=1e-3, nbumps=nbumps, bump_amp_max=0.2, bump_sigma_min=1.0, bump_sigma_max=3.0,
                     seed=seed, manning_n=0.03, inflow_q=q, tailwater=tail, ic_flat_wse=flat)
    return spec, flume_mesh(spec)


def perturbed_state(sol0, seed=0, amp_h=0.05, amp_v=0.3):
    rng = np.random.default_rng(seed)
    s = sol0.copy()
    n = len(s) // 3
    s[0::3] += amp_h * rng.uniform(-1, 1, n)
    s[1::3] += amp_v * rng.uniform(-1, 1, n)
    s[2::3] += amp_v * rng.uniform(-1, 1, n)
    so = sol0.copy()
    so[0::3] += amp_h * rng.uniform(-1, 1, n)
    so[1::3] += amp_v * rng.uniform(-1, 1, n)
    so[2::3] += amp_v * rng.uniform(-1, 1, n)
    return s, so


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)
