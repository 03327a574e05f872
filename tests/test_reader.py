"""Host input path on the CPU: the product's Matrix Market reader and CSR builder
(generate_CSC semantics, code/read_input.cu:11-171) against the step-by-step
restatement in oracle/reader.py, plus the generators and the id remapping.  No GPU."""
import numpy as np
import pytest

from oracle import reader as R


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _check(g, ref):
    row, col, w = g.csr()
    np.testing.assert_array_equal(row, ref["row"])
    np.testing.assert_array_equal(col, ref["col"])
    np.testing.assert_array_equal(w, ref["w"])
    assert g.nv == ref["nv"] and g.ne == ref["ne"] and g.start_p == ref["start_p"]


# general integer matrix: duplicates (same weight), a self-loop, isolated vertices 3 and 8,
# comment lines, entries out of order
GENERAL = """%%MatrixMarket matrix coordinate integer general
% a comment
%another % comment
9 9 9
2 1 5
1 2 5
4 5 7
5 4 7
2 1 5
6 6 3
7 9 11
9 7 11
1 7 2
"""

# symmetric header: lower triangle only, real values (truncated toward zero)
SYMREAL = """%%MatrixMarket matrix coordinate real symmetric
6 6 5
2 1 2.7
3 2 -1.5
5 4 9.99
6 1 4.0
6 5 1e1
"""

COMPLEX = """%%MatrixMarket matrix coordinate complex general
4 4 4
1 2 3.5 -2.0
2 1 3.5 -2.0
3 4 7.25 1.0
4 3 7.25 1.0
"""


# Hermitian header: read as symmetric (read_input.cu:31), real part of each value
HERMITIAN = """%%MatrixMarket matrix coordinate complex Hermitian
5 5 4
2 1 3.5 -2.0
3 1 -7.75 1.0
4 3 12.0 0.5
5 2 1.25 0.0
"""


@pytest.mark.parametrize("text", [GENERAL, SYMREAL, COMPLEX, HERMITIAN], ids=["general", "symreal", "complex", "hermitian"])
@pytest.mark.parametrize("flags", ["SYM", "NONE", "HEADER", "SYM|CC"])
@pytest.mark.parametrize("source", [0, 2, 5])
def test_reader_matches_restatement(mg, tmp_path, text, flags, source):
    path = _write(tmp_path, "g.mtx", text)
    nv, hsym, nflag, entries = R.parse_mtx(text)
    if source >= nv:
        pytest.skip("source outside this matrix")
    f = {"SYM": mg.READ_SYM, "NONE": 0, "HEADER": mg.READ_HEADER, "SYM|CC": mg.READ_SYM | mg.READ_CC}[flags]
    sym = hsym if flags == "HEADER" else ("SYM" in flags)
    cc = "CC" in flags
    try:
        ref = R.generate_csc(nv, entries, sym, cc, source)
        ok = ref["nv"] > 0 and all(c < ref["nv"] for c in ref["col"]) and \
            all(r <= ref["ne"] for r in ref["row"])
    except IndexError:
        ok = False
    if not ok:
        # entry 0 of the sorted order names a vertex no other entry does: the reference
        # writes past its arrays; the product refuses the input
        with pytest.raises(mg.MGError):
            mg.Graph.load_mtx(path, source=source, flags=f)
        return
    g = mg.Graph.load_mtx(path, source=source, flags=f)
    _check(g, ref)
    # the compacted -> original id map is the inverse of valid[]
    orig = g.orig_ids()
    valid = ref["valid"]
    for c, o in enumerate(orig):
        assert valid[o] == c and valid[o + 1] == c + 1
    assert g.remap(source) == ref["start_p"]


def test_entry0_quirk(mg, tmp_path):
    """read_input.cu:106-114 never marks the endpoints of sorted entry 0 valid: vertex 1
    (0-based 0) below has no other entry, so its row is filed under the next valid id and
    the merged row holds the same column twice, as the reference's CSR does"""
    text = "%%MatrixMarket matrix coordinate integer general\n5 5 3\n1 3 4\n2 3 6\n3 2 6\n"
    path = _write(tmp_path, "q.mtx", text)
    nv, _, _, entries = R.parse_mtx(text)
    ref = R.generate_csc(nv, entries, False, False, 0)
    g = mg.Graph.load_mtx(path, source=0, flags=0)
    _check(g, ref)
    row, col, _ = g.csr()
    assert list(row) == [0, 2, 3] and list(col) == [1, 1, 0]


def test_entry0_out_of_range_rejected(mg, tmp_path):
    text = "%%MatrixMarket matrix coordinate integer general\n6 6 3\n1 6 4\n3 4 6\n4 3 6\n"
    path = _write(tmp_path, "bad0.mtx", text)
    with pytest.raises(mg.MGError) as ei:
        mg.Graph.load_mtx(path, source=0, flags=0)
    assert ei.value.code == mg.ERR_RANGE


def test_pattern_weights(mg, tmp_path):
    text = "%%MatrixMarket matrix coordinate pattern symmetric\n% c\n5 5 4\n2 1\n3 2\n4 3\n5 1\n"
    path = _write(tmp_path, "p.mtx", text)
    nv, _, nflag, entries = R.parse_mtx(text)
    assert nflag == 0
    g1 = mg.Graph.load_mtx(path, flags=mg.READ_SYM, weight_seed=1)
    g2 = mg.Graph.load_mtx(path, flags=mg.READ_SYM, weight_seed=1)
    g3 = mg.Graph.load_mtx(path, flags=mg.READ_SYM, weight_seed=99)
    r1, c1, w1 = g1.csr()
    _, _, w2 = g2.csr()
    _, _, w3 = g3.csr()
    np.testing.assert_array_equal(w1, w2)                      # seeded, deterministic
    assert w1.min() >= 1 and w1.max() <= 64                     # rand()%64+1 range, :77
    assert not np.array_equal(w1, w3)
    # structure is the restatement's; both directions of an entry carry one weight
    ref = R.generate_csc(nv, [(s, d, 0) for s, d, _ in entries], True, False, 0)
    np.testing.assert_array_equal(r1, ref["row"])
    np.testing.assert_array_equal(c1, ref["col"])
    for u in range(g1.nv):
        for p in range(r1[u], r1[u + 1]):
            v = c1[p]
            q = r1[v] + np.searchsorted(c1[r1[v]:r1[v + 1]], u)
            assert w1[q] == w1[p]


def test_reader_errors(mg, tmp_path):
    bad = _write(tmp_path, "oor.mtx", "%%MatrixMarket matrix coordinate integer general\n3 3 2\n1 2 1\n4 1 1\n")
    with pytest.raises(mg.MGError) as ei:
        mg.Graph.load_mtx(bad)
    assert ei.value.code == mg.ERR_RANGE and "out of range" in str(ei.value)
    with pytest.raises(mg.MGError) as ei:
        mg.Graph.load_mtx(str(tmp_path / "missing.mtx"))
    assert ei.value.code == mg.ERR_IO
    nohdr = _write(tmp_path, "nohdr.mtx", "3 3 1\n1 2 1\n")
    with pytest.raises(mg.MGError):
        mg.Graph.load_mtx(nohdr)
    short = _write(tmp_path, "short.mtx", "%%MatrixMarket matrix coordinate integer general\n3 3 2\n1 2 1\n")
    with pytest.raises(mg.MGError):
        mg.Graph.load_mtx(short)
    with pytest.raises(mg.MGError):
        mg.Graph.load_mtx(_write(tmp_path, "g.mtx", GENERAL), source=9)


def test_from_coo_matches_restatement(mg):
    rng = np.random.default_rng(4)
    nv = 40
    s = rng.integers(0, nv, 150)
    d = rng.integers(0, nv, 150)
    v = rng.integers(1, 65, 150)
    entries = [(int(a), int(b), int(c)) for a, b, c in zip(s, d, v)]
    for flags, sym, cc in ((mg.READ_SYM, True, False), (mg.READ_SYM | mg.READ_CC, True, True)):
        g = mg.Graph.from_coo(nv, s, d, v, source=3, flags=flags)
        _check(g, R.generate_csc(nv, entries, sym, cc, 3))


def test_rmat_generator(mg):
    g1 = mg.Graph.rmat(10, 8, seed=3, weight_seed=5)
    g2 = mg.Graph.rmat(10, 8, seed=3, weight_seed=5)
    r1, c1, w1 = g1.csr()
    r2, c2, w2 = g2.csr()
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(c1, c2)
    np.testing.assert_array_equal(w1, w2)
    src = np.repeat(np.arange(g1.nv), np.diff(r1))
    assert not np.any(src == c1)                                   # self-loops dropped
    assert w1.min() >= 1 and w1.max() <= 64
    key = set(zip(src.tolist(), c1.tolist(), w1.tolist()))
    assert all((b, a, x) in key for a, b, x in key)                # symmetric, one weight per pair
    for u in range(g1.nv):                                          # rows sorted, no duplicates
        assert np.all(np.diff(c1[r1[u]:r1[u + 1]]) > 0)
    assert g1.nv <= 1 << 10 and g1.ne <= 2 * 8 * (1 << 10)
    gd = mg.Graph.rmat(10, 8, seed=3, weight_seed=5, flags=0)     # directed read keeps one direction
    assert gd.ne < g1.ne


def test_grid_generator_and_roundtrip(mg, tmp_path):
    W, H = 7, 5
    g = mg.Graph.grid(W, H)
    row, col, w = g.csr()
    assert g.nv == W * H and g.ne == 2 * ((W - 1) * H + W * (H - 1))
    for y in range(H):
        for x in range(W):
            i = y * W + x
            nb = sorted([j for j in (i - W, i - 1, i + 1, i + W)
                         if 0 <= j < W * H and (j // W == y or j % W == x)])
            assert list(col[row[i]:row[i + 1]]) == nb
    path = str(tmp_path / "grid.mtx")
    g.write_mtx(path)
    g2 = mg.Graph.load_mtx(path, flags=mg.READ_HEADER)
    r2, c2, w2 = g2.csr()
    np.testing.assert_array_equal(row, r2)
    np.testing.assert_array_equal(col, c2)
    np.testing.assert_array_equal(w, w2)


def test_union(mg):
    a = mg.Graph.grid(3, 3)
    b = mg.Graph.rmat(6, 4, seed=2)
    u = mg.Graph.union([a, b])
    ra, ca, wa = a.csr()
    rb, cb, wb = b.csr()
    ru, cu, wu = u.csr()
    assert u.nv == a.nv + b.nv and u.ne == a.ne + b.ne
    np.testing.assert_array_equal(ru[:a.nv + 1], ra)
    np.testing.assert_array_equal(ru[a.nv:], rb + a.ne)
    np.testing.assert_array_equal(cu, np.concatenate([ca, cb + a.nv]))
    np.testing.assert_array_equal(wu, np.concatenate([wa, wb]))


def test_max_degree_vertex(mg):
    g = mg.Graph.rmat(9, 8, seed=4)
    row, _, _ = g.csr()
    deg = np.diff(row)
    assert deg[g.max_degree_vertex()] == deg.max()


# ---- pins of the restatement that do not go through it ---------------------------
# Hand-computed CSRs of the fixtures above (0-based ids, isolated ids compacted away,
# rows sorted by column, real values truncated toward zero).
HAND = {
    # GENERAL, symmetrized: vertices 0 1 3 4 5 6 8 -> 0..6 (2 and 7 isolated); the
    # self-loop on 5 is kept, the three copies of {0,1} collapse to one edge each way
    ("general", "SYM"): dict(row=[0, 2, 3, 4, 5, 6, 8, 9], col=[1, 5, 0, 3, 2, 4, 0, 6, 5],
                             w=[5, 2, 5, 7, 7, 3, 2, 11, 11], src5=4),
    # GENERAL as given: 0->6 has no reverse
    ("general", "NONE"): dict(row=[0, 2, 3, 4, 5, 6, 7, 8], col=[1, 5, 0, 3, 2, 4, 6, 5],
                              w=[5, 2, 5, 7, 7, 3, 11, 11], src5=4),
    # SYMREAL: 2.7 -> 2, -1.5 -> -1, 9.99 -> 9, 4.0 -> 4, 1e1 -> 10
    ("symreal", "SYM"): dict(row=[0, 2, 4, 5, 6, 8, 10], col=[1, 5, 0, 2, 1, 4, 3, 5, 0, 4],
                             w=[2, 4, 2, -1, -1, 9, 9, 10, 4, 10], src5=5),
    # SYMREAL with the CC rule (src < dst only): rows 2 and 5 are empty but kept
    ("symreal", "SYM|CC"): dict(row=[0, 2, 3, 3, 4, 5, 5], col=[1, 5, 2, 4, 5], w=[2, 4, -1, 9, 10], src5=5),
    # HERMITIAN, real parts: 3.5 -> 3, -7.75 -> -7, 12.0 -> 12, 1.25 -> 1
    ("hermitian", "SYM"): dict(row=[0, 2, 4, 6, 7, 8], col=[1, 2, 0, 4, 0, 3, 2, 1], w=[3, -7, 3, 1, -7, 12, 12, 1],
                               src5=None),
}


@pytest.mark.parametrize("key", list(HAND), ids=["-".join(k) for k in HAND])
def test_reader_hand_computed(mg, tmp_path, key):
    text = {"general": GENERAL, "symreal": SYMREAL, "hermitian": HERMITIAN}[key[0]]
    f = {"SYM": mg.READ_SYM, "NONE": 0, "SYM|CC": mg.READ_SYM | mg.READ_CC}[key[1]]
    h = HAND[key]
    path = _write(tmp_path, "h.mtx", text)
    nv, _, _, entries = R.parse_mtx(text)
    for g in (mg.Graph.load_mtx(path, source=0, flags=f),):
        row, col, w = g.csr()
        assert list(row) == h["row"] and list(col) == h["col"] and list(w) == h["w"]
    ref = R.generate_csc(nv, entries, "SYM" in key[1], "CC" in key[1], 0)
    assert list(ref["row"]) == h["row"] and list(ref["col"]) == h["col"] and list(ref["w"]) == h["w"]
    if h["src5"] is not None:
        assert mg.Graph.load_mtx(path, source=5, flags=f).start_p == h["src5"]


def _scipy_csr(path, sym, cc):
    """the reader's rules rebuilt from scipy.io.mmread's COO: symmetrize, drop repeated
    (src, dst) pairs (inputs here give repeats one value), keep src < dst for CC,
    compact the ids that occur in some entry, CSR with rows sorted by column"""
    import scipy.io

    A = scipy.io.mmread(path).tocoo()
    r, c = A.row.astype(np.int64), A.col.astype(np.int64)
    v = np.trunc(np.real(A.data)).astype(np.int64)
    n = A.shape[0]
    if sym:
        r, c, v = np.concatenate([r, c]), np.concatenate([c, r]), np.concatenate([v, v])
    used = np.zeros(n, bool)
    used[r] = True
    used[c] = True
    key = r * n + c
    key, first = np.unique(key, return_index=True)
    r, c, v = key // n, key % n, v[first]
    if cc:
        keep = r < c
        r, c, v = r[keep], c[keep], v[keep]
    newid = np.cumsum(used) - 1
    nv = int(used.sum())
    r, c = newid[r], newid[c]
    row = np.zeros(nv + 1, np.int64)
    np.add.at(row, r + 1, 1)
    return np.cumsum(row), c, v, nv


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("flags", ["SYM", "SYM|CC", "NONE"])
def test_reader_vs_scipy_mmread(mg, tmp_path, seed, flags):
    """random general integer matrices with repeated pairs, self-loops and isolated ids,
    read by the product and by scipy.io.mmread plus the rules above"""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 60))
    m = int(rng.integers(1, 4 * n))
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    if seed % 3 == 0:  # some repeated pairs
        s = np.concatenate([s, s[: m // 3]])
        d = np.concatenate([d, d[: m // 3]])
    val = (np.minimum(s, d) * 131 + np.maximum(s, d) * 7) % 64 + 1  # one value per unordered pair
    sym, cc = "SYM" in flags, "CC" in flags
    # inputs whose first sorted entry names an id no other entry names take the
    # reference's entry-0 quirk (tested above): leave them out here
    pairs = sorted(set(zip(s.tolist(), d.tolist())) | (set(zip(d.tolist(), s.tolist())) if sym else set()))
    u0, v0 = pairs[0]
    rest = pairs[1:]
    if not any(u0 in p for p in rest) or not any(v0 in p for p in rest):
        pytest.skip("entry-0 quirk input")
    text = "%%MatrixMarket matrix coordinate integer general\n" + f"{n} {n} {len(s)}\n" + \
        "".join(f"{a + 1} {b + 1} {x}\n" for a, b, x in zip(s, d, val))
    path = _write(tmp_path, "r.mtx", text)
    f = (mg.READ_SYM if sym else 0) | (mg.READ_CC if cc else 0)
    g = mg.Graph.load_mtx(path, source=0, flags=f)
    row, col, w = g.csr()
    er, ec, ev, env = _scipy_csr(path, sym, cc)
    assert g.nv == env
    np.testing.assert_array_equal(row, er)
    np.testing.assert_array_equal(col, ec)
    np.testing.assert_array_equal(w, ev)
    # and the step-by-step restatement agrees with scipy too
    nv, _, _, entries = R.parse_mtx(text)
    ref = R.generate_csc(nv, entries, sym, cc, 0)
    np.testing.assert_array_equal(ref["row"], er)
    np.testing.assert_array_equal(ref["col"], ec)
