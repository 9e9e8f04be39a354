"""Parity helpers shared by the GPU tests: compare GPU hits with the oracle's, classify mismatches."""
import numpy as np

MISS = 0xFFFFFFFF


def classify(gpu, ref, brute=None):
    """Return counts of agreeing rays and of each mismatch class.

    gpu / ref: tuples (prim uint32, t, b0, b1) as numpy arrays.
    Classes: 'tie' (same t, different triangle), 'ref_false_cull' (the brute
    force agrees with the GPU), 'unexplained'.
    """
    gp, gt, gb0, gb1 = gpu
    rp, rt, rb0, rb1 = ref
    same = (gp == rp) & (gt.view(np.uint32) == rt.view(np.uint32))
    tie = (~same) & (gt == rt) & (gp != rp)
    other = ~same & ~tie
    false_cull = np.zeros_like(other)
    if brute is not None and other.any():
        bp, bt = brute
        false_cull = other & (bp == gp) & (bt.view(np.uint32) == gt.view(np.uint32))
    unexplained = other & ~false_cull
    bary_bad = same & ((gb0.view(np.uint32) != rb0.view(np.uint32)) |
                       (gb1.view(np.uint32) != rb1.view(np.uint32)))
    return {
        "n": int(len(gp)), "agree": int(same.sum()), "tie": int(tie.sum()),
        "ref_false_cull": int(false_cull.sum()), "unexplained": int(unexplained.sum()),
        "bary_mismatch": int(bary_bad.sum()), "hits": int((gp != MISS).sum()),
        "unexplained_idx": np.where(unexplained)[0][:10].tolist(),
    }


def max_rel_t_err(gpu_t, ref_t, mask):
    g = gpu_t[mask].astype(np.float64)
    r = ref_t[mask].astype(np.float64)
    if len(g) == 0:
        return 0.0
    return float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-30)))
