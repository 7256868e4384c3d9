"""Writes the golden fixtures in this folder.  Run from the repository root:

    python tests/golden/make_fixtures.py

Inputs are the reference's own geometry data files (read as text):
  Single_phase/geo_cavity.dat -- 50^3 cavity mask written by
      Single_phase/flow_domain_geo_generation_2D.py, one value per line, x fastest
  docs/geo.dat                -- 48^3 all-fluid mask
They are stored bit-packed, in file order.  The compaction checksums are the
SHA-256 of the fluid list and neighbour table, plus the face records, that
the plain-loop restatement (oracle.compaction.compact_loops) produces for the
cavity mask with the example_cavity.py boundary (velocity on x-right,
SP:405-407).

pack32_compaction.npz holds a seeded 32^3 overlapping-sphere pack (this
package's generator, seed 7, radii 3-6, porosity 0.35, periodic in y and z)
bit-packed in C order, with the fluid list and the full neighbour table from
compact_loops for pressure faces on x: the checksums for a porous medium, where
bounce-back and periodic wrap both occur.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)


def pack_file(src, n, dst):
    vals = np.loadtxt(src)
    assert vals.size == n
    bits = (vals > 0).astype(np.uint8)
    np.savez_compressed(dst, bits=np.packbits(bits), n=n, n_solid=int(bits.sum()))


def main():
    pack_file(os.path.join(REF, "Single_phase", "geo_cavity.dat"), 50 ** 3, os.path.join(HERE, "geo_cavity.npz"))
    pack_file(os.path.join(REF, "docs", "geo.dat"), 48 ** 3, os.path.join(HERE, "geo_docs48.npz"))
    from oracle.compaction import compact_loops
    d = np.load(os.path.join(HERE, "geo_cavity.npz"))
    bits = np.unpackbits(d["bits"])[: int(d["n"])]
    solid = bits.reshape((50, 50, 50), order="F")
    modes = (0, 2, 0, 0, 0, 0)
    cells, rec, nbr = compact_loops(solid, modes)
    np.savez_compressed(os.path.join(HERE, "cavity_compaction.npz"), n_cells=cells.size, modes=np.array(modes),
                        face_rec=rec.astype(np.uint8),
                        sha_cells=hashlib.sha256(cells.astype("<i4").tobytes()).hexdigest(),
                        sha_nbr=hashlib.sha256(nbr.astype("<i4").tobytes()).hexdigest())
    print("wrote fixtures:", cells.size, "fluid cells,", int((rec > 0).sum()), "face-owned")
    from taichi_lbm3d_amd.geometry import sphere_pack
    pack = sphere_pack((32, 32, 32), porosity=0.35, rmin=3.0, rmax=6.0, seed=7, periodic=(False, True, True))
    modes = (1, 1, 0, 0, 0, 0)
    cells, rec, nbr = compact_loops(pack, modes)
    np.savez_compressed(os.path.join(HERE, "pack32_compaction.npz"), bits=np.packbits(pack.reshape(-1)),
                        modes=np.array(modes), cells=cells.astype(np.int32), nbr=nbr.astype(np.int32),
                        face_rec=rec.astype(np.uint8))
    print("pack32:", cells.size, "fluid cells")


if __name__ == "__main__":
    main()
