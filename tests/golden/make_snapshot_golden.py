"""Writes the ASCII snapshot fixtures snapshot_2d.dat and snapshot_3d.dat.

The format is restated here in plain Python from the reference's writers
(FileHeader::writeAscii, src/class.h:54-57: "%e\\n" time, "%d\\n" N;
RealPtcl::writeAscii, src/class.h:129-135: 12 tab-separated columns,
"%ld" for id and tag, "%.16lf" for the rest; 2D writes z = vz = 0 and the
sound speed, 3D writes snds = 0). Python's % formatting rounds correctly,
as glibc's printf does, so the text is what the reference would print.

Run from the repository root: python tests/golden/make_snapshot_golden.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))

# (id, tag, mass, pos, vel, dens, snds, pres): values chosen to exercise
# rounding in the 16th decimal, negative numbers, large and tiny magnitudes
PARTICLES = [
    (0, 0, 0.002000376, (0.0063, 0.0063, 0.0063), (0.0, 0.0, 0.0), 1000.1234567890123, 62.6418, 17.25),
    (7, 1, 1.0 / 3.0, (-0.0189, 0.5, 2.9937), (1.5e-3, -2.25, 1.0e-17), 999.0, 62.5, 0.0),
    (12345678901, 0, 2.0 ** -30, (3.9999999999999996, 0.1 + 0.2, 1.0e-9), (-8.86, 0.0, -0.3333333333333333),
     1012.5, 63.0, 123456.78901234567),
]
TIME = 0.41666666666666669


def line(dim, p):
    pid, tag, mass, pos, vel, dens, snds, pres = p
    if dim == 2:
        cols = [mass, pos[0], pos[1], 0.0, vel[0], vel[1], 0.0, dens, snds, pres]
    else:
        cols = [mass, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], dens, 0.0, pres]
    return "%d\t%d\t" % (pid, tag) + "\t".join("%.16f" % c for c in cols) + "\n"


def text(dim):
    return "%e\n%d\n" % (TIME, len(PARTICLES)) + "".join(line(dim, p) for p in PARTICLES)


if __name__ == "__main__":
    for dim in (2, 3):
        with open(os.path.join(HERE, f"snapshot_{dim}d.dat"), "w") as f:
            f.write(text(dim))
