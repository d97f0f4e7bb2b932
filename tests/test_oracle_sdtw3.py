"""The three-move oracle (reference built without -DNO_REF_DEL,
src/include/DTW.cu:23-27): D[r][c] = cost + min(D[r][c-1], D[r-1][c],
D[r-1][c-1]), free start, column -1 = +inf.

Pinned by an independent definition -- the minimum over every monotone path
(steps down, diagonal, right) from row 0 to the last row -- enumerated in
pure Python on tiny inputs, by the two-move score as an upper bound, the
Q = 1 closed form, exact substrings, and per-cell numpy float32/float16
restatements on small inputs."""
import functools

import numpy as np

import oracle as O


def _paths_min(q, s):
    """min over monotone paths; a path starts anywhere in row 0, moves
    (r+1, c), (r+1, c+1) or (r, c+1), ends anywhere in the last row, and
    costs the sum of (s_c - q_r)^2 over the cells it visits."""
    Q, N = len(q), len(s)

    @functools.lru_cache(maxsize=None)
    def best_from(r, c):  # cheapest completion of a path that is at (r, c)
        here = (s[c] - q[r]) ** 2
        if r == Q - 1:
            # on the last row it may stop here or keep moving right
            return here + (min(0.0, best_from(r, c + 1)) if c + 1 < N else 0.0)
        opts = [best_from(r + 1, c)]
        if c + 1 < N:
            opts += [best_from(r + 1, c + 1), best_from(r, c + 1)]
        return here + min(opts)

    return min(best_from(0, c) for c in range(N))


def test_three_moves_equal_path_enumeration():
    rng = np.random.default_rng(7)
    for Q, N in ((1, 1), (1, 5), (2, 3), (3, 3), (4, 5), (5, 4), (3, 7)):
        for _ in range(5):
            q = rng.standard_normal(Q).round(3)
            s = rng.standard_normal(N).round(3)
            sc, am = O.sdtw3_f64(q, s)
            assert abs(sc - _paths_min(tuple(q), tuple(s))) < 1e-12, (Q, N)
            # two-move paths are a subset
            assert sc <= O.sdtw_f64(q, s)[0] + 1e-12


def test_three_moves_closed_forms():
    rng = np.random.default_rng(8)
    s = rng.standard_normal(50)
    q1 = np.array([0.3])
    assert O.sdtw3_f64(q1, s)[0] == np.min((s - 0.3) ** 2)
    assert O.sdtw3_f64(s[10:30], s)[0] == 0.0
    # a horizontal run: one query value against a constant stretch of the
    # reference costs nothing more with right moves; with two moves it does
    s2 = np.concatenate([rng.standard_normal(5) + 5, np.zeros(6), rng.standard_normal(5) + 5])
    q2 = np.zeros(2)
    assert O.sdtw3_f64(q2, s2)[0] == 0.0


def _f32_restated(q, s):
    Q, N = len(q), len(s)
    D = np.full((Q + 1, N + 1), np.float32(np.inf), np.float32)
    D[0, 1:] = 0
    for r in range(Q):
        for c in range(N):
            t = np.float32(s[c] - q[r])
            m = min(D[r + 1, c], D[r, c + 1], D[r, c])
            D[r + 1, c + 1] = np.float32(np.float32(t * t) + m)
    return D[Q, 1:]


def _f16_restated(qh, sh):
    q = qh.view(np.float16)
    s = sh.view(np.float16)
    Q, N = len(q), len(s)
    D = np.full((Q + 1, N + 1), np.float16(np.inf), np.float16)
    D[0, 1:] = 0
    for r in range(Q):
        for c in range(N):
            t = np.float16(s[c] - q[r])
            m = min(D[r + 1, c], D[r, c + 1], D[r, c])
            D[r + 1, c + 1] = np.float16(np.float16(t * t) + m)
    return D[Q, 1:]


def test_three_moves_f32_f16_per_cell():
    rng = np.random.default_rng(9)
    for Q, N in ((6, 9), (17, 23), (30, 40)):
        q = rng.standard_normal(Q).astype(np.float32)
        s = rng.standard_normal(N).astype(np.float32)
        row = _f32_restated(q, s)
        sc, am = O.sdtw3_f32(q, s)
        assert sc == row.min() and am == int(np.argmin(row))
        qh, sh = O.f32_to_f16(q), O.f32_to_f16(s)
        row16 = _f16_restated(qh, sh)
        sc16, am16 = O.sdtw3_f16(qh, sh)
        assert np.uint16(sc16).view(np.float16) == row16.min() and am16 == int(np.argmin(row16))


def test_batch3_matches_single():
    rng = np.random.default_rng(10)
    ref = rng.standard_normal((2, 300)).astype(np.float32)
    reads = [rng.standard_normal(n).astype(np.float32) for n in (1, 7, 64, 129)]
    sc, am = O.batch3_f32(reads, ref, n_threads=2)
    refh = np.stack([O.f32_to_f16(ref[k]) for k in range(2)])
    qh = [O.f32_to_f16(r) for r in reads]
    sch, amh = O.batch3_f16(qh, refh, n_threads=2)
    for i, r in enumerate(reads):
        for k in range(2):
            assert (sc[i, k], am[i, k]) == O.sdtw3_f32(r, ref[k])
            assert (sch[i, k], amh[i, k]) == O.sdtw3_f16(qh[i], refh[k])
