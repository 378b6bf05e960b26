"""gridstates.bin / gridinput.bin: reading snapshots and writing committor inputs.

Formats (src/io.c):
  gridstates.bin  records appended at every grid output (write_ising_grids,
                  src/io.c:67-130): int32 L, int32 ngrids, int32 isweep, then
                  the L*L*ngrids spins of all grids as one bit stream (grid-
                  major, site-major, bit k of byte b is spin 8b + k, 1 = +1),
                  padded to whole bytes.
  gridinput.bin   one grid (read_input_grid, src/io.c:4-65): int32 L, then
                  ceil(L*L / 8) bytes of the same bit layout.

The reference's viewer writes gridinput.bin from the snapshot on screen
(tools/vis_gridstates.py:109-114, assuming L*L divisible by 8); `extract`
does the same without a display, for any L:

  python -m gpu_2dising_amd.gridfiles info gridstates.bin
  python -m gpu_2dising_amd.gridfiles extract gridstates.bin SNAPSHOT REPLICA [-o gridinput.bin]
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def read_gridstates(path: str):
    """All records of a gridstates.bin file: list of (isweep, int8 grids
    [ngrids, L, L] with values +-1)."""
    data = np.fromfile(path, dtype=np.uint8)
    out, pos = [], 0
    while pos < data.size:
        if pos + 12 > data.size:
            raise ValueError(f"{path}: truncated record header at byte {pos}")
        L, ngrids, isweep = (int(v) for v in data[pos:pos + 12].view(np.int32))
        if L <= 0 or ngrids <= 0:
            raise ValueError(f"{path}: bad record header (L={L}, ngrids={ngrids}) at byte {pos}")
        nbits = L * L * ngrids
        nbytes = (nbits + 7) // 8
        body = data[pos + 12:pos + 12 + nbytes]
        if body.size != nbytes:
            raise ValueError(f"{path}: truncated record at byte {pos}")
        bits = np.unpackbits(body, bitorder="little")[:nbits]
        out.append((isweep, (2 * bits.astype(np.int8) - 1).reshape(ngrids, L, L)))
        pos += 12 + nbytes
    return out


def write_gridinput(path: str, grid) -> None:
    """One +-1 grid [L, L] as gridinput.bin."""
    g = np.asarray(grid)
    if g.ndim != 2 or g.shape[0] != g.shape[1]:
        raise ValueError("grid must be square")
    L = g.shape[0]
    bits = np.packbits((g.reshape(-1) > 0).astype(np.uint8), bitorder="little")
    with open(path, "wb") as f:
        f.write(np.int32(L).tobytes())
        f.write(bits.tobytes())


def extract(gridstates: str, snapshot: int, replica: int, out: str = "gridinput.bin") -> np.ndarray:
    """Writes replica `replica` of record `snapshot` of a gridstates file to
    `out` (gridinput.bin) and returns the grid."""
    recs = read_gridstates(gridstates)
    if not -len(recs) <= snapshot < len(recs):
        raise IndexError(f"snapshot {snapshot} out of range ({len(recs)} records)")
    _, grids = recs[snapshot]
    if not 0 <= replica < grids.shape[0]:
        raise IndexError(f"replica {replica} out of range ({grids.shape[0]} grids)")
    write_gridinput(out, grids[replica])
    return grids[replica]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m gpu_2dising_amd.gridfiles", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("info", help="list the records of a gridstates.bin file")
    p.add_argument("gridstates")
    p = sub.add_parser("extract", help="write one replica of one record as gridinput.bin")
    p.add_argument("gridstates")
    p.add_argument("snapshot", type=int)
    p.add_argument("replica", type=int)
    p.add_argument("-o", "--out", default="gridinput.bin")
    a = ap.parse_args(argv)
    if a.cmd == "info":
        for k, (isweep, grids) in enumerate(read_gridstates(a.gridstates)):
            m = grids.reshape(grids.shape[0], -1).mean(1)
            print(f"record {k}: sweep {isweep}, {grids.shape[0]} grids of L={grids.shape[1]}, "
                  f"magnetisation min {m.min():+.4f} max {m.max():+.4f}")
    else:
        g = extract(a.gridstates, a.snapshot, a.replica, a.out)
        print(f"Grid snapshot written to {a.out} (L={g.shape[0]}, magnetisation {g.mean():+.4f})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
