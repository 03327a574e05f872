"""Pure-Python restatement of the reference's input path, generate_CSC
(code/read_input.cu:11-171), step by step and with the reference's own arrays
(temp_v, loc, valid), for small inputs.

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the product's Matrix
Market reader and CSR builder (multigraph_amd/csrc/host_graph.cpp).  The product
never imports it.

The one choice the reference leaves open is which of several duplicate (src, dst)
entries keeps its value: read_input.cu:97 sorts with the unstable C qsort.  Here,
as in the product, the sort is stable, so the first entry in input order wins
(parity for duplicate-weight inputs is unpinned by the reference).  Pattern files
get rand()-based weights in the reference (read_input.cu:21,77); here the caller
passes the weights the product drew, so only the structure is restated.
"""
from __future__ import annotations

import numpy as np

INF = 0x64646464


class ReaderError(Exception):
    pass


def parse_mtx(text: str):
    """read_input.cu:23-96: returns (nv, sflag_header, nflag, entries) with 0-based
    (src, dst, val) entries in file order; val is None for pattern files."""
    lines = text.split("\n")
    header = lines[0]
    sflag = "symmetric" in header or "Hermitian" in header          # :31
    if "pattern" in header:                                            # :33-35
        nflag = 0
    elif "complex" in header:
        nflag = -1
    else:
        nflag = 1
    i = 1
    while i < len(lines) and "%" in lines[i]:                           # :46-51
        i += 1
    toks = " ".join(lines[i:]).split()
    nv, _cols, nnz = int(toks[0]), int(toks[1]), int(toks[2])         # :57
    per = 2 if nflag == 0 else (3 if nflag == 1 else 4)
    entries = []
    k = 3
    for _ in range(nnz):
        s, d = int(toks[k]) - 1, int(toks[k + 1]) - 1                   # :67-68
        if s < 0 or s >= nv or d < 0 or d >= nv:                        # :70-73
            raise ReaderError(f"A vertex id is out of range {s} {d}")
        val = None
        if nflag != 0:
            val = int(np.float32(float(toks[k + 2])))                   # :78-86, (int)float
        entries.append((s, d, val))
        k += per
    return nv, sflag, nflag, entries


def generate_csc(nv, entries, sym: bool, cc: bool, start: int, weights=None):
    """read_input.cu:88-171 on 0-based entries.  Returns dict(row, col, w, nv, ne,
    start_p, valid).  `weights` replaces the values entry by entry (pattern input)."""
    temp = []
    for i, (s, d, v) in enumerate(entries):
        if weights is not None:
            v = int(weights[i])
        temp.append((s, d, v))                                          # :66-87
        if sym:
            temp.append((d, s, v))                                      # :88-95
    ne = len(temp)
    temp.sort(key=lambda e: (e[0], e[1]))                              # :97 (stable here)
    loc = [0] * (ne + 1)
    valid = [0] * (nv + 1)
    if ne > 0:
        loc[0] = 1                                                      # :105
    for i in range(1, ne):                                              # :106-114
        valid[temp[i][0]] = 1
        valid[temp[i][1]] = 1
        loc[i] = 0 if (temp[i][0] == temp[i - 1][0] and temp[i][1] == temp[i - 1][1]) else 1
        if cc and temp[i][0] >= temp[i][1]:
            loc[i] = 0
    for i in range(1, nv + 1):                                          # :115-119 exclusive scan
        valid[i] += valid[i - 1]
    for i in range(nv, 0, -1):
        valid[i] = valid[i - 1]
    valid[0] = 0
    for i in range(1, ne + 1):                                          # :120-124
        loc[i] += loc[i - 1]
    for i in range(ne, 0, -1):
        loc[i] = loc[i - 1]
    loc[0] = 0
    start_p = valid[start]                                              # :130
    out = [None] * ne
    for i in range(ne):                                                 # :135-139
        s, d, v = temp[i]
        out[loc[i]] = (valid[s], valid[d], v)
    cnv, cne = valid[nv], loc[ne]                                       # :146
    out = out[:cne]
    row = [0] * (cnv + 1)                                               # :150-171
    col = [0] * cne
    w = [0] * cne
    for t, (s, d, v) in enumerate(out):
        col[t] = d
        w[t] = v
        if 1 + s <= cnv:
            row[1 + s] = t + 1
    for i in range(1, cnv):
        if row[i] == 0:
            row[i] = row[i - 1]
    if cnv >= 0:
        row[cnv] = cne
    return dict(row=np.array(row, np.int32), col=np.array(col, np.int32), w=np.array(w, np.int32),
                nv=cnv, ne=cne, start_p=start_p, valid=np.array(valid, np.int64))
