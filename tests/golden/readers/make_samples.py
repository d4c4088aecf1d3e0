"""Write the sample structure files of tests/golden/readers/ (one per format
of the reference's readers, atomicStructure.f90:143-475). They are data:
a perovskite unit cell (lattice slightly sheared, atoms jittered) and the
methylammonium molecule, in the units each format uses.

    python tests/golden/readers/make_samples.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "..", "high-dimensional-neural-network-potential_amd"))
import workloads as wl  # noqa: E402

ANG2BOHR = 1.8897259886
SYM = {1: "H", 6: "C", 7: "N", 53: "I", 82: "Pb"}


def main():
    s = wl.perovskite(1, 1, 1, seed=3)
    lat = np.array([[11.9, 0.0, 0.0], [0.35, 11.7, 0.0], [-0.2, 0.15, 12.1]])   # row k = vector k, bohr
    pos = s.pos @ np.linalg.solve(np.diag([11.9] * 3), lat)                      # same fractional coords
    lat_a, pos_a = lat / ANG2BOHR, pos / ANG2BOHR
    frac = pos @ np.linalg.inv(lat)
    m = wl.methylammonium()

    with open(os.path.join(HERE, "geometry.in"), "w") as f:          # FHI-aims, periodic
        f.write("# perovskite unit cell\n")
        for v in lat_a:
            f.write("lattice_vector %16.10f %16.10f %16.10f\n" % tuple(v))
        for k, (z, p) in enumerate(zip(s.Z, pos_a)):
            f.write("atom %16.10f %16.10f %16.10f %s\n" % (*p, SYM[int(z)]))
            if k == 2:
                f.write("  initial_moment 0.0\n")
        f.write("  atom 0.0 0.0 0.0 H\n")          # indented: not an atom line for the reader
    with open(os.path.join(HERE, "molecule.in"), "w") as f:          # FHI-aims, no lattice
        for z, p in zip(m.Z, m.pos / ANG2BOHR):
            f.write("atom %14.8f %14.8f %14.8f %s\n" % (*p, SYM[int(z)]))
    for name, unit, key in (("cell_ang.ascii", 1 / ANG2BOHR, "angstroem"), ("cell_bohr.ascii", 1.0, "")):
        L, P = lat * unit, pos * unit
        with open(os.path.join(HERE, name), "w") as f:              # v_sim ascii: lower-triangular lattice
            f.write(f"{s.nat} {key}\n")
            f.write("%18.12f %18.12f %18.12f\n" % (L[0, 0], L[1, 0], L[1, 1]))
            f.write("%18.12f %18.12f %18.12f\n" % (L[2, 0], L[2, 1], L[2, 2]))
            for z, p in zip(s.Z, P):
                f.write("%18.12f %18.12f %18.12f %s\n" % (*p, SYM[int(z)]))
    order = [82, 53, 7, 6, 1]
    with open(os.path.join(HERE, "POSCAR"), "w") as f:               # VASP, direct coordinates
        f.write("perovskite unit cell\n   2.0\n")                    # the scale is read, not applied
        for v in lat_a:
            f.write("  %16.10f %16.10f %16.10f\n" % tuple(v))
        f.write("   " + "   ".join(SYM[z] for z in order) + "\n")
        f.write("   " + "   ".join(str(int(np.sum(s.Z == z))) for z in order) + "\n")
        f.write("Direct\n")
        for z in order:
            for p in frac[s.Z == z]:
                f.write("  %16.12f %16.12f %16.12f\n" % tuple(p))
    rng = np.random.default_rng(4)
    with open(os.path.join(HERE, "input.data"), "w") as f:           # RuNNer data, bohr and Ha
        for blk, (Z, P, L) in enumerate(((s.Z, pos, lat), (m.Z, m.pos, None))):
            f.write("begin\n")
            f.write(f"comment structure {blk + 1}\n")
            if L is not None:
                for v in L:
                    f.write("lattice %18.10f %18.10f %18.10f\n" % tuple(v))
            q = rng.normal(0, 0.2, len(Z))
            F = rng.normal(0, 0.05, (len(Z), 3))
            for z, p, qq, ff in zip(Z, P, q, F):
                f.write("atom %16.10f %16.10f %16.10f %-2s %12.8f %12.8f %14.10f %14.10f %14.10f\n"
                        % (*p, SYM[int(z)], qq, 0.0, *ff))
            f.write("energy %20.10f\n" % (-40.5 - 3.25 * blk))
            f.write("charge %d\n" % (1 - blk))
            f.write("end\n")
    big = wl.perovskite(2, 2, 1, seed=6)                               # 48 atoms: the readers grow their arrays
    with open(os.path.join(HERE, "big.data"), "w") as f:
        for blk in range(2):
            f.write("  begin\n")
            for v in big.lat:
                f.write("  lattice %18.10f %18.10f %18.10f\n" % tuple(v))
            for z, p in zip(big.Z, big.pos + blk * 0.1):
                f.write("  atom %16.10f %16.10f %16.10f %s %8.4f 0.0 %8.4f %8.4f %8.4f\n"
                        % (*p, SYM[int(z)], 0.01 * blk, 0.001, -0.002, 0.003))
            f.write("  energy -1.5d2\n  charge 0\n  end\n")


if __name__ == "__main__":
    main()
