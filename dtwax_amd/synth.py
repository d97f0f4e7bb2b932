"""Seeded synthetic workloads for the sDTW classifier (reference, pore model,
target and non-target reads).

Generator (one value per k-mer, both strands):
  * 6-mer model: 4096 level means uniform in [60, 130) pA, sd 1.5.
  * reference: uniform random ACGT sequence of `ref_bp` bases; its squiggle
    is the z-scored per-k-mer level of both strands (dtwax_squiggle_build).
  * target read: a random strand and start position of the reference; each
    k-mer dwells a geometric number of samples (mean 9, support >= 1) and
    every sample gets N(0, 2 pA) noise.
  * non-target read: the same process on an unrelated random sequence.
Raw reads are in pA; normalisation (median/MAD or mean/std) happens on the
GPU.  Everything is reproducible from the seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BASES = np.frombuffer(b"ACGT", dtype=np.uint8)


def random_seq(rng, n):
    return BASES[rng.integers(0, 4, n)].tobytes().decode()


def kmer_codes(seq: str, k: int = 6) -> np.ndarray:
    a = np.frombuffer(seq.encode(), np.uint8)
    lut = np.full(256, -1, np.int64)
    for i, b in enumerate(b"ACGT"):
        lut[b] = i
    v = lut[a]
    n = len(seq) - k + 1
    code = np.zeros(n, np.int64)
    bad = np.zeros(n, bool)
    for j in range(k):
        code = code * 4 + np.maximum(v[j:j + n], 0)
        bad |= v[j:j + n] < 0
    code[bad] = -1
    return code


def revcomp(seq: str) -> str:
    return seq.translate(str.maketrans("ACGT", "TGCA"))[::-1]


def make_model(rng, k=6, lo=60.0, hi=130.0, sd=1.5):
    means = rng.uniform(lo, hi, 4 ** k).astype(np.float32)
    sds = np.full(4 ** k, sd, np.float32)
    return means, sds


def simulate_signal(rng, levels, n_samples, dwell_mean=9.0, noise_sd=2.0):
    """levels: per-k-mer pA levels (long enough); returns n_samples float32."""
    need = int(n_samples / dwell_mean * 1.5) + 64
    lv = levels[:need]
    while True:
        dw = rng.geometric(1.0 / dwell_mean, size=lv.size)
        if dw.sum() >= n_samples or lv.size >= levels.size:
            break
        need *= 2
        lv = levels[:need]
    sig = np.repeat(lv, dw)[:n_samples]
    if sig.size < n_samples:  # ran out of reference: wrap (only for tiny references)
        sig = np.resize(sig, n_samples)
    return (sig + rng.normal(0.0, noise_sd, n_samples)).astype(np.float32)


@dataclass
class Workload:
    seq: str
    model_means: np.ndarray
    ref_levels: np.ndarray           # (2, n_cols) z-scored squiggle (fwd, rev)
    reads: list = field(default_factory=list)     # raw pA reads (float32)
    is_target: np.ndarray = None                  # bool per read
    origin: np.ndarray = None                     # (strand, kmer start) per target read, -1 else


def medmad_z(x):
    """(x - median) / (1.4826 * MAD) in float64, returned as float32."""
    x = np.asarray(x, np.float64)
    med = np.median(x)
    mad = np.median(np.abs(x - med))
    return ((x - med) / (1.4826 * mad)).astype(np.float32)


def build_reference(rng, ref_bp, model_means, k=6, ref_norm="medmad"):
    """ref_norm: "meanstd" -- the reference squiggle arithmetic of
    load_reference.hpp:33-108 (dtwax_squiggle_build); "medmad" -- the same
    robust estimator the reads get, per strand.  With the uniform level model
    above, 1.4826*MAD is 1.28x the standard deviation, so mixing estimators
    would put targets 28% out of scale."""
    from . import squiggle_build
    seq = random_seq(rng, ref_bp)
    if ref_norm == "meanstd":
        levels, skipped = squiggle_build(seq, model_means, None, k)
        assert skipped == 0
    else:
        levels = np.stack([medmad_z(model_means[kmer_codes(s, k)]) for s in (seq, revcomp(seq))])
    return seq, levels


def make_reads(rng, seq, model_means, lengths, target_frac, k=6):
    """Reads of the given raw lengths; the first round(target_frac*n) positions
    are chosen at random to be targets."""
    n = len(lengths)
    n_t = int(round(target_frac * n))
    is_t = np.zeros(n, bool)
    is_t[rng.permutation(n)[:n_t]] = True
    fwd_codes = kmer_codes(seq, k)
    rev_codes = kmer_codes(revcomp(seq), k)
    reads, origin = [], np.full((n, 2), -1, np.int64)
    for i, L in enumerate(lengths):
        if is_t[i]:
            strand = int(rng.integers(0, 2))
            codes = fwd_codes if strand == 0 else rev_codes
            span = int(L / 9.0 * 1.5) + 64
            start = int(rng.integers(0, max(1, codes.size - span)))
            lv = model_means[codes[start:start + span]]
            origin[i] = (strand, start)
        else:
            other = random_seq(rng, int(L / 9.0 * 1.5) + 64 + k)
            lv = model_means[kmer_codes(other, k)]
        reads.append(simulate_signal(rng, lv, int(L)))
    return reads, is_t, origin


def config_workload(seed, ref_bp, n_reads, raw_len, target_frac=0.5, len_range=None,
                    ref_norm="medmad"):
    """len_range=(lo, hi): uniform integer raw lengths; else all raw_len."""
    rng = np.random.default_rng(seed)
    means, _ = make_model(rng)
    seq, levels = build_reference(rng, ref_bp, means, ref_norm=ref_norm)
    if len_range is not None:
        lengths = rng.integers(len_range[0], len_range[1] + 1, n_reads)
    else:
        lengths = np.full(n_reads, raw_len, np.int64)
    reads, is_t, origin = make_reads(rng, seq, means, lengths, target_frac)
    return Workload(seq, means, levels, reads, is_t, origin)


# Named configurations (sizes as stated for each workload).
CONFIGS = {
    # 10 kb reference, 64 reads x 2000 samples, 50% target, FP32, CPU-runnable
    "c1_fp32_10kb": dict(seed=1, ref_bp=10_000, n_reads=64, raw_len=2000, trim=0),
    # 48.5 kb reference, 4096 reads x 3000 raw -> trim 1000 -> 2000, 50% target, FP16
    "c2_fp16_48kb": dict(seed=2, ref_bp=48_500, n_reads=4096, raw_len=3000, trim=1000),
}


def trim_reads(reads, trim):
    return [r[trim:] for r in reads] if trim else reads
