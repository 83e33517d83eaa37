"""Seeded R-MAT edge lists with PaRMAT's construction.

The reference builds its R-MAT datasets with PaRMAT in sorted, undirected,
loop-free, duplicate-free mode (scripts/rmat_preparation.sh:39-46:
``-sorted -noEdgeToSelf -noDuplicateEdges -undirected -threads 8``) and turns
each edge list into a symmetric all-ones matrix with rmat_ones.  PaRMAT seeds
its generators from std::random_device and time(0), so no run can be
reproduced bit for bit; this module keeps every step of its construction and
draws from a numpy seed instead, so the graphs have the same distribution:

* the vertex range is inclusive, [0, nVertices - 1] (PaRMAT/src/PaRMAT.cpp:144);
* squares split at the integer midpoint (x0 + x1) >> 1, which both halves
  keep as a boundary vertex (PaRMAT/src/Square.cpp:32-50);
* a split hands its edge count to the four parts by the shares a, b, c (+ a
  small noise) truncated to integers, the last part taking the remainder;
  on the diagonal of an undirected graph the b part is dropped and its edges
  go to the c part (PaRMAT/src/utils.cpp:63-99);
* the sorted generator splits every square until a column of squares holds at
  most ``capacity`` edges and there are at least ``threads`` columns, unless
  a square gets too dense (3 * edges >= area) (GraphGen_sorted.cpp:87-110,
  utils.cpp:27-59,140-146);
* the two coordinates of an edge are drawn independently, each by halving a
  floating-point interval with the cut line a + c (row) or a + b (column)
  until it is narrower than 1 and rounding its midpoint (utils.cpp:101-121);
* an undirected graph keeps row <= column inside diagonal squares, loops are
  redrawn, and duplicates inside a square are redrawn until the square holds
  distinct edges (utils.cpp:149-172, GraphGen_sorted.cpp:17-71);
* every column of squares is sorted by (row, column) and the columns are
  written in order, one "row<TAB>column" line per edge (Edge.cpp:23-41).

Two choices stand in for what PaRMAT leaves to its host: ``capacity``
(PaRMAT derives it from the machine's RAM; the default never splits for
capacity, as on any host with more than a few GB) and the six split noises,
which PaRMAT draws from rand() after srand(time(0)) on every split, so all
splits within one second share them; one draw per graph stands for that.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .csr import Csr

RAND_MAX = 2**31 - 1  # glibc rand() and std::uniform_int_distribution<int>'s max


@dataclass
class Square:
    """PaRMAT/src/Square.hpp: an inclusive block [x0, x1] x [y0, y1] of the
    adjacency matrix, its edge budget and its (horizontal, vertical) index."""
    x0: int
    x1: int
    y0: int
    y1: int
    n: int
    h: int = 0
    v: int = 0


def split_noise(rng) -> np.ndarray:
    """ShatterSquare's six noises (rand() - RAND_MAX/2) / (RAND_MAX * 100):
    the three cut-line noises, then the three edge-share noises."""
    r = rng.integers(0, RAND_MAX, 6, endpoint=True)
    return (r - RAND_MAX // 2) / (RAND_MAX * 100.0)


def shatter(sq: Square, a: float, b: float, c: float, noise, directed: bool) -> list[Square]:
    """The parts of one split, in PaRMAT's order (a, b, c, d; the b part
    dropped on the diagonal of an undirected graph)."""
    sa, sb, sc = a + noise[3], b + noise[4], c + noise[5]
    xm, ym = (sq.x0 + sq.x1) >> 1, (sq.y0 + sq.y1) >> 1
    parts = []
    for p, share in enumerate((sa, sb, sc, 1.0 - sa - sb - sc)):
        right, low = p in (1, 3), p in (2, 3)
        parts.append(Square(xm if right else sq.x0, sq.x1 if right else xm,
                            ym if low else sq.y0, sq.y1 if low else ym,
                            int(share * float(sq.n)), 2 * sq.h + right, 2 * sq.v + low))
    parts[3].n = sq.n - parts[0].n - parts[1].n - parts[2].n
    if not directed and sq.x1 == sq.y1:
        parts[2].n += parts[1].n
        del parts[1]
    return parts


def _columns_within(squares: list[Square], capacity: int) -> bool:
    """Eligible_RNG_Squares: no run of squares with one x1 holds more than
    ``capacity`` edges."""
    run, x1 = 0, 0
    for sq in squares:
        run = run + sq.n if sq.x1 == x1 else sq.n
        x1 = sq.x1
        if run > capacity:
            return False
    return True


def _num_columns(squares: list[Square]) -> int:
    cols, x1 = 0, 0
    for sq in squares:
        if sq.x1 != x1:
            x1 = sq.x1
            cols += 1
    return cols


def _too_dense(squares: list[Square]) -> bool:
    return any(3 * sq.n >= (sq.x1 - sq.x0) * (sq.y1 - sq.y0) for sq in squares)


def plan_squares(n_vertices: int, n_edges: int, a: float, b: float, c: float, noise, threads: int = 8,
                 capacity: int | None = None, directed: bool = False) -> list[Square]:
    """The squares the sorted generator draws from, ordered by (h, v)."""
    if n_vertices < 2 or n_edges == 0 or n_edges >= n_vertices * n_vertices:
        raise ValueError("number of edges or vertices out of range")
    if n_vertices > 1 << 30:
        raise ValueError("more than 2^30 vertices are not supported")
    if n_edges < 10000:
        threads = 1
    cap = capacity if capacity is not None else n_edges
    squares = [Square(0, n_vertices - 1, 0, n_vertices - 1, n_edges)]

    def split_all(sqs):
        out = [p for sq in sqs for p in shatter(sq, a, b, c, noise, directed)]
        return sorted(out, key=lambda s: (s.h, s.v))

    while not _columns_within(squares, cap):
        squares = split_all(squares)
    while _num_columns(squares) < threads and not _too_dense(squares):
        squares = split_all(squares)
    return squares


def below_count(cut: float) -> int:
    """Number of draws i in [0, RAND_MAX] with i / RAND_MAX < cut in double
    arithmetic; i / RAND_MAX grows with i, so those are [0, count)."""
    k = min(max(int(cut * RAND_MAX), 0), RAND_MAX + 1)
    while k <= RAND_MAX and k / RAND_MAX < cut:
        k += 1
    while k > 0 and (k - 1) / RAND_MAX >= cut:
        k -= 1
    return k


def halvings(lo: int, hi: int) -> int:
    """How often genEdgeIndex_FP halves [lo, hi]: while the width is at
    least 1, i.e. bit_length(hi - lo) times."""
    return (hi - lo).bit_length()


def index_from_halvings(lo: int, hi: int, j: np.ndarray) -> np.ndarray:
    """genEdgeIndex_FP's result for the halving choices ``j`` (k bits, most
    significant first, a set bit keeping the upper half).  Below 2^30 every
    interval end is exact in double precision, s = lo + w j / 2^k with
    w = hi - lo, so floor((s + e) / 2 + 0.5) is the integer
    lo + floor((w (2j + 1) + 2^k) / 2^(k + 1))."""
    if hi >= 1 << 30:
        raise ValueError("vertex indices from 2^30 up are not supported")
    w, k = hi - lo, halvings(lo, hi)
    return lo + ((w * (2 * j + 1) + (1 << k)) >> (k + 1))


def axis_index(rng, lo: int, hi: int, cut: float, count: int) -> np.ndarray:
    """genEdgeIndex_FP (utils.cpp:101-121) for ``count`` edges: each halving
    keeps the lower half when a uniform draw, an int in [0, RAND_MAX] over
    RAND_MAX, is below ``cut``."""
    thr = 2 * below_count(cut)   # u32 < 2 * count  <=>  (u32 >> 1) < count
    j = np.zeros(count, np.int64)
    for _ in range(halvings(lo, hi)):
        j <<= 1
        if thr < 1 << 32:
            j |= rng.integers(0, 1 << 32, count, dtype=np.uint32) >= thr
    return index_from_halvings(lo, hi, j)


def square_edges(rng, sq: Square, a: float, b: float, c: float, directed=False, self_loops=False,
                 duplicates=False) -> tuple[np.ndarray, np.ndarray]:
    """generate_edges plus the per-square duplicate redraw: ``sq.n`` edges
    (distinct unless ``duplicates``), sorted by (row, column) unless
    duplicates are allowed."""
    keep_upper = not (directed or sq.h < sq.v)
    span = sq.y1 + 1

    def draw(count):
        hs, vs, got = [], [], 0
        while got < count:
            m = (count - got) + (count - got) // 4 + 16
            h = axis_index(rng, sq.x0, sq.x1, a + c, m)
            v = axis_index(rng, sq.y0, sq.y1, a + b, m)
            ok = np.ones(m, bool)
            if keep_upper:
                ok &= h <= v
            if not self_loops:
                ok &= h != v
            h, v = h[ok][:count - got], v[ok][:count - got]
            hs.append(h)
            vs.append(v)
            got += len(h)
        return np.concatenate(hs) if hs else np.empty(0, np.int64), np.concatenate(vs) if vs else np.empty(0, np.int64)

    h, v = draw(sq.n)
    if duplicates:
        return h, v
    keys = np.unique(h * span + v)
    for _ in range(100000):
        if len(keys) == sq.n:
            return keys // span, keys % span
        h2, v2 = draw(sq.n - len(keys))
        keys = np.unique(np.concatenate([keys, h2 * span + v2]))
    raise RuntimeError("square too dense for distinct edges")


def parmat_edges(n_vertices: int, n_edges: int, a: float = 0.45, b: float = 0.22, c: float = 0.22, seed: int = 0,
                 threads: int = 8, capacity: int | None = None, directed: bool = False, self_loops: bool = False,
                 duplicates: bool = False, noise=None) -> tuple[np.ndarray, np.ndarray]:
    """(row, column) arrays of the sorted generator's output lines, in file
    order.  Defaults are the reference's dataset flags; a, b, c default to
    PaRMAT's own (0.45, 0.22, 0.22)."""
    rng = np.random.default_rng(seed)
    noise = split_noise(rng) if noise is None else np.asarray(noise, np.float64)
    squares = plan_squares(n_vertices, n_edges, a, b, c, noise, threads, capacity, directed)
    rows, cols = [], []
    i = 0
    while i < len(squares):  # one column of squares (equal x0) per worker
        j = i
        while j < len(squares) and squares[j].x0 == squares[i].x0:
            j += 1
        hs, vs = zip(*(square_edges(rng, sq, a, b, c, directed, self_loops, duplicates) for sq in squares[i:j]))
        keys = np.sort(np.concatenate(hs) * n_vertices + np.concatenate(vs))
        rows.append(keys // n_vertices)
        cols.append(keys % n_vertices)
        i = j
    return np.concatenate(rows), np.concatenate(cols)


def write_txt(path: str, rows: np.ndarray, cols: np.ndarray) -> None:
    """PaRMAT's text output: one "row<TAB>column" line per edge."""
    with open(path, "w") as f:
        f.write("".join(f"{r}\t{c}\n" for r, c in zip(rows.tolist(), cols.tolist())))


def parmat_symmetric(scale: int, ef: int, abc=(0.57, 0.19, 0.19), seed: int = 0, threads: int = 8) -> Csr:
    """The matrix rmat_ones makes of a PaRMAT dataset "<name>_<scale>_<ef>":
    every edge and its mirror, all values 1.0, sorted by (row, column),
    duplicates kept (src/mm/mm_utils.c:757-847)."""
    n = 1 << scale
    h, v = parmat_edges(n, ef * n, *abc, seed=seed, threads=threads)
    keys = np.sort(np.concatenate([h * n + v, v * n + h]))
    rows, cols = keys // n, keys % n
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return Csr(n, n, rp, cols.astype(np.int32), np.ones(len(rows), np.float64))
