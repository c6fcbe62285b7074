"""Shared comparisons for the parity tests."""
import numpy as np

MISS = 0xFFFFFFFF


def tied(oracle_out):
    """Rays whose closest hit is tied within 1e-6 in t with another triangle."""
    h = oracle_out["hits"]
    return (h["inst"] != MISS) & ((oracle_out["second_t"] - h["t"]) <= 1e-6)


def compare_hits(gpu_hits, oracle_out):
    """(non-tied id mismatches, t mismatches among equal ids, tied count)."""
    h = oracle_out["hits"]
    t = tied(oracle_out)
    same_id = (gpu_hits["inst"] == h["inst"]) & (gpu_hits["tri"] == h["tri"])
    id_bad = int((~same_id & ~t).sum())
    t_bad = int((same_id & (gpu_hits["t"] != h["t"])).sum())
    return id_bad, t_bad, int(t.sum())


def pixel_rel_err(a_xyz, b_xyz):
    """Per-pixel max-abs XYZ difference relative to the pixel's largest |XYZ|."""
    scale = np.maximum(np.abs(b_xyz).max(axis=-1), 1e-30)
    return np.abs(a_xyz - b_xyz).max(axis=-1) / scale
