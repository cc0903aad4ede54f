#!/usr/bin/env python3
"""Generate ``bg_tables.inc``: the NR LDPC base graphs in the row/column/V form of
3GPP TS 38.212 Tables 5.3.2-2 (BG1) and 5.3.2-3 (BG2).

The only machine-readable copy of the two tables available offline is the reference's
``lib/phy/upper/channel_coding/ldpc/ldpc_luts_impl.cpp:434-4365`` (dense 46x68 matrices,
one per lifting-set index, ``NO_EDGE`` marking absent edges).  This script reads that
file as text, extracts the nonzero entries and writes them in the standard's sparse
layout: for each base-graph row i, the sorted column indices j and the eight shift
values V_{i,j} (one per lifting-set index iLS = 0..7).  The row adjacency lists in the
same file (``:4367-4503``) are parsed too and must equal the sorted column lists; the
script refuses to write the output otherwise.

Usage: python gen_bg_tables.py <reference_root> <output.inc>
"""
import re
import sys
from pathlib import Path

BG_DIMS = {1: (46, 68), 2: (42, 52)}
MAX_M, MAX_N = 46, 68


def _tokens(text):
    return re.findall(r"NO_EDGE|\d+", text)


def _section(src, start_pat, end_pat):
    s = src.index(start_pat)
    e = src.index(end_pat, s)
    return src[s:e]


def parse(ref_root):
    path = Path(ref_root) / "lib/phy/upper/channel_coding/ldpc/ldpc_luts_impl.cpp"
    src = path.read_text()
    out = {}
    for bg, nxt in ((1, "BG2_matrices = {"), (2, "BG1_adjacency_matrix = {")):
        body = _section(src, f"BG{bg}_matrices = {{", nxt)
        # drop comments ("lifting size index k" would add stray integers)
        body = re.sub(r"//[^\n]*", "", body)
        row_txts = re.findall(r"\{([^{}]*)\}", body.split("=", 1)[1])
        M, N = BG_DIMS[bg]
        if len(row_txts) != 8 * M:
            raise SystemExit(f"BG{bg}: expected {8 * M} rows, found {len(row_txts)}")
        grid = []
        for t in row_txts:
            toks = _tokens(t)
            if len(toks) != N:
                raise SystemExit(f"BG{bg}: row with {len(toks)} entries, expected {N}")
            grid.append(toks)
        rows = []
        for i in range(M):
            cols = []
            for j in range(N):
                vals = []
                for ils in range(8):
                    t = grid[ils * M + i][j]
                    vals.append(None if t == "NO_EDGE" else int(t))
                present = [v is not None for v in vals]
                if any(present) and not all(present):
                    raise SystemExit(f"BG{bg} ({i},{j}): edge present in only some lifting sets")
                if all(present):
                    cols.append((j, vals))
            rows.append(cols)
        out[bg] = rows
    # cross-check the adjacency lists
    for bg, nxt in ((1, "BG2_adjacency_matrix = {"), (2, "BG_matrix_t ocudu::ldpc::get_graph")):
        body = _section(src, f"BG{bg}_adjacency_matrix = {{", nxt)
        body = re.sub(r"//[^\n]*", "", body)
        inner = body.split("=", 1)[1]
        row_txts = re.findall(r"\{([^{}]*)\}", inner)
        M, _ = BG_DIMS[bg]
        if len(row_txts) < M:
            raise SystemExit(f"BG{bg}: adjacency rows {len(row_txts)} < {M}")
        for i in range(M):
            adj = [int(t) for t in _tokens(row_txts[i]) if t != "NO_EDGE"]
            cols = [j for j, _ in out[bg][i]]
            if adj != cols:
                raise SystemExit(f"BG{bg} row {i}: adjacency {adj} != sorted columns {cols}")
    return out


def emit(tables, dst):
    lines = [
        "// Generated by gen_bg_tables.py -- do not edit.",
        "// NR LDPC base graphs, 3GPP TS 38.212 Table 5.3.2-2 (BG1) and Table 5.3.2-3 (BG2),",
        "// in the standard's sparse form: for row i, the column indices j (ascending) and the",
        "// shift coefficients V_{i,j} for lifting-set index iLS = 0..7.  The lifted shift for",
        "// lifting size Z is V_{i,j} mod Z.",
        "#pragma once",
        "#include <stdint.h>",
        "#ifdef __cplusplus",
        "#define LH_TBL static constexpr",
        "#else",
        "#define LH_TBL static const",
        "#endif",
        "",
    ]
    for bg, rows in tables.items():
        M = len(rows)
        nnz = sum(len(r) for r in rows)
        starts = [0]
        for r in rows:
            starts.append(starts[-1] + len(r))
        cols = [j for r in rows for j, _ in r]
        lines.append(f"#define LH_BG{bg}_NNZ {nnz}")
        lines.append(f"LH_TBL uint16_t lh_bg{bg}_row_start[{M + 1}] = {{{', '.join(map(str, starts))}}};")
        lines.append(f"LH_TBL uint8_t lh_bg{bg}_col[{nnz}] = {{{', '.join(map(str, cols))}}};")
        lines.append(f"LH_TBL uint16_t lh_bg{bg}_v[8][{nnz}] = {{")
        for ils in range(8):
            vals = [v[ils] for r in rows for _, v in r]
            lines.append("  {" + ", ".join(map(str, vals)) + "},")
        lines.append("};")
        lines.append("")
    Path(dst).write_text("\n".join(lines))


if __name__ == "__main__":
    if len(sys.argv) != 3:
        raise SystemExit(__doc__)
    emit(parse(sys.argv[1]), sys.argv[2])
