"""Shared helpers of the parity tests: scene generation and tolerance checks."""
import numpy as np

from tsr_amd import SyntheticScene

PARAMS = (2.0, 100.0, 25.0, 1.0, 2.0)   # App.cfg:63-73


def scene(config, scale=1.0, seed=0):
    s = SyntheticScene(config, seed=seed, resolution_scale=scale)
    return s.samples, s.cameras


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, np.abs(a - b) / den, np.where(a == b, 0.0, np.inf))
    return r


def occupancy_rel_err(occ_gpu, oracle_field, eps=1e-6):
    """|occ_gpu - occ_oracle| / (pS + pE) per leaf (Occupancy.cpp:570-587), with the
    -EPSILON leaves required to match exactly."""
    f = oracle_field
    L = len(f["occupancy"])
    E, S = f["kernel_sums"][:L], f["kernel_sums"][L:]
    M = f["cone_count"]
    pr = f["priors"]
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = np.where(np.isfinite(pr), pr * E / M + (1 - pr) * S / M, 0.0)
    o = f["occupancy"]
    neg = o < 0
    err = np.zeros(L)
    mask = ~neg
    err[mask] = np.abs(occ_gpu[mask] - o[mask]) / np.maximum(scale[mask], 1e-300)
    err[neg] = np.where(occ_gpu[neg] == o[neg], 0.0, np.abs(occ_gpu[neg] - o[neg]) / np.maximum(scale[neg], 1e-300))
    flips = int(np.sum((occ_gpu > 0) != (o > 0)))
    return err, flips
