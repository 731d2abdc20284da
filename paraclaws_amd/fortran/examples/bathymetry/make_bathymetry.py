#!/usr/bin/env python3
"""Write bathymetry.bin for the grid of ./control.in: the seeded sum of eight
Gaussians (seed 3; amplitudes 0-0.8, widths 0.05-0.3, centres in the domain)
over the frame cells i = -1 .. nx+2, j = -1 .. ny+2, as doubles with i
fastest (the Fortran array b(-1:nx+2, -1:ny+2))."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..", "..", "..")))

from paraclaws_amd import ic  # noqa: E402


def main(path="control.in", out="bathymetry.bin", seed=3):
    vals = [ln.split()[0] for ln in open(path) if ln.strip()]
    nx, ny = int(vals[18]), int(vals[19])
    xi = (float(vals[20]), float(vals[21]))
    eta = (float(vals[22]), float(vals[23]))
    ic.gaussian_bathymetry(nx, ny, seed=seed, xi=xi, eta=eta).tofile(out)


if __name__ == "__main__":
    main(*sys.argv[1:])
