"""Shared helpers for the test suite (fixture loading, python-form reference)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_model_means(path=None):
    """6-mer model file -> (dict kmer->mean as Python float, means[4096] float32
    indexed by 2-bit code A=0 C=1 G=2 T=3, first base most significant)."""
    path = path or os.path.join(GOLDEN, "dna_kmer_model.txt")
    d = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) >= 2:
                d[p[0]] = float(p[1])
    codes = {"A": 0, "C": 1, "G": 2, "T": 3}
    arr = np.zeros(4096, np.float32)
    for k, v in d.items():
        c = 0
        for ch in k:
            c = c * 4 + codes[ch]
        arr[c] = np.float32(v)
    return d, arr


def fasta_python(path=None):
    """scripts/sdtw.py:60-63 get_fasta: every line after the first, joined."""
    path = path or os.path.join(GOLDEN, "reference.fasta")
    with open(path) as f:
        return "".join(f.read().split("\n")[1:])


def rev_comp_python(bases):
    """scripts/sdtw.py:66-68."""
    return bases.replace("A", "t").replace("T", "a").replace("G", "c").replace("C", "g").upper()[::-1]


def golden_reference_f64():
    """fwd[:48128] | rev[:48128], each python-z-normalised (scripts/sdtw.py:101-116, 234)."""
    import oracle as O
    model, _ = load_model_means()
    fa = fasta_python()

    def sig(seq):
        v = np.array([model[seq[i:i + 6]] for i in range(len(seq) - 6 + 1)], np.float64)
        return O.python_znorm(v[:48128])

    return np.concatenate([sig(fa), sig(rev_comp_python(fa))])


# ---- the reference's launch layouts (src/main.cu:92-109, 165-179) ----------
REF_TILE, REF_SEG, REF_WARP = 832, 26, 32


def tile_reference(words):
    """Flat 32-bit words (REF_LEN,) -> the reference's tiled layout
    h[l*832 + i*32 + j] = flat[l*832 + j*26 + i] over REF_LEN // 832 tiles;
    the tail beyond them is never copied (left 0)."""
    words = np.ascontiguousarray(words)
    n_tiles = words.shape[0] // REF_TILE
    c = np.arange(n_tiles * REF_TILE)
    l, rem = c // REF_TILE, c % REF_TILE
    out = np.zeros_like(words)
    out[l * REF_TILE + (rem % REF_SEG) * REF_WARP + rem // REF_SEG] = words[c]
    return out


def tiled_index(cols, ref_len):
    """Index of each column in the reference's tiled last-row layout (the
    last tile at ref_len - 832, src/include/DTW.cu:451, 546)."""
    cols = np.asarray(cols)
    n_tiles = ref_len // REF_TILE
    t, w = cols // REF_TILE, cols % REF_TILE
    base = np.where(t < n_tiles - 1, t * REF_TILE, ref_len - REF_TILE)
    return base + w // REF_SEG + REF_WARP * (w % REF_SEG)


def half2_words(fwd, rev):
    """Two float arrays (binary16-representable or not: rounded RN) -> uint32
    words {fwd, rev} as __floats2half2_rn packs them."""
    a = np.asarray(fwd, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    b = np.asarray(rev, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    return a | (b << 16)


def query_words_h2(qs):
    """(n, 2048) float queries -> n*2048 + 32 half2 words {q, q} (zero pad)."""
    w = half2_words(qs.ravel(), qs.ravel())
    return np.concatenate([w, np.zeros(32, np.uint32)])


def words_to_h2_floats(words):
    """uint32 half2 words -> (n, 2) float32 (fwd, rev)."""
    return np.ascontiguousarray(words, np.uint32).view(np.float16).reshape(-1, 2).astype(np.float32)


def norm_chunks_medmad(raw, chunk_len):
    """The Read Until engine's median/MAD normalisation: the first chunk's
    statistics applied to every chunk, y = (float)((x - centre) * inv) in
    binary64 (inv = 1 / (1.4826 * MAD), 0 for a constant chunk)."""
    import oracle as O
    first = raw[:chunk_len]
    y0, med, mad = O.norm_medmad(first)
    scale = 1.4826 * mad
    inv = 1.0 / scale if scale > 0 else 0.0
    y = ((np.asarray(raw, np.float64) - med) * inv).astype(np.float32)
    assert np.array_equal(y[:chunk_len], y0)
    return y
