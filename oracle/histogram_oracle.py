"""CPU restatement of the reference's 256-bin histogram and KL divergence.

TEST INFRASTRUCTURE ONLY (the oracle): used by tests/ as the checker of the
HIP path, never by the product.

* quantize + histogram256 + merge_histogram256 (runner.hpp:4425-4539,
  launched by Runner::launch_histogram256[_with_mask], runner.hpp:5849-5890):
  bin = floor((v - v_min) * step_inverse) mod 256 in FP32, step_inverse =
  float(256 / (double(v_max) - double(v_min))); the FP32 -> u32 conversion
  rounds down and saturates (negative and NaN -> 0).  With a mask, only
  elements whose mask is exactly 1 count.
* calculate_kl_divergence (runner.hpp:4803-4815) with the KLDivergence
  functor (runner.hpp:746-762): sum over bins with p_b != 0 of
  p_b / n_p * log(p_b / max(q_b, q_lower_bound) * n_q / n_p), terms in FP32.
"""
import numpy as np


def histogram256(v, v_min, v_max, mask=None):
    v = np.asarray(v, np.float32)
    step_inverse = np.float32(256.0 / (float(v_max) - float(v_min)))
    with np.errstate(invalid="ignore", over="ignore"):
        x = (v - np.float32(v_min)) * step_inverse  # FP32 arithmetic
    q = np.zeros(x.shape, np.uint64)
    ok = x >= 0  # False for NaN and negatives
    big = ok & (x >= np.float32(4294967296.0))
    mid = ok & ~big
    q[mid] = np.floor(x[mid].astype(np.float64)).astype(np.uint64)
    q[big] = 0xFFFFFFFF
    bins = (q & 0xFF).astype(np.int64)
    if mask is not None:
        bins = bins[np.asarray(mask, np.float32) == np.float32(1.0)]
    return np.bincount(bins, minlength=256).astype(np.uint32)


def kl_divergence(hp, hq, n_p, n_q, q_lower_bound, num_bins):
    p = np.asarray(hp[:num_bins], np.float32)
    q = np.asarray(hq[:num_bins], np.float32)
    n_p_inv = np.float32(1.0) / np.float32(n_p)
    ratio = np.float32(n_q) / np.float32(n_p)
    nz = p != 0
    with np.errstate(divide="ignore"):
        terms = p[nz] * n_p_inv * np.log(p[nz] / np.maximum(q[nz], np.float32(q_lower_bound)) * ratio)
    return float(np.sum(terms.astype(np.float64)))
