"""Copies two data files of the reference's own tests into tests/golden/ (run
once, in a tree that has the reference; the copies are what the tests read):

  * silicon.gap: tests/stopping/1/gap_files/silicon.gap, a potential file with
    distance_2b, angle_3b, soap_turbo (compressed) and core_pot blocks;
  * gap_files/alphas_2b.dat, gap_files/silicon.xml.sparseX..._9741,
    gap_files/alphas_3b.dat, gap_files/silicon.xml.sparseX..._9742 and
    gap_files/corepot.dat: the distance_2b and angle_3b sparse sets and the
    core_pot table that silicon.gap names (paths relative to tests/golden, as the .gap gives
    them relative to the run directory);
  * gap_files/compress.dat and gap_files/alphas_mb.dat: the soap_turbo block's
    compression file and its alphas (its desc_sparse file, ..._9743, is not in
    the reference tree);
  * xrd_mad_frame0.xyz: the first frame of
    tests/regression/cases/xrd_mad/expected/trajectory_out.xyz, written by the
    reference's write_extxyz.
"""
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(REF, "tests/stopping/1/gap_files/silicon.gap")) as f:
    gap = f.read()
with open(os.path.join(HERE, "silicon.gap"), "w") as f:
    f.write(gap)

os.makedirs(os.path.join(HERE, "gap_files"), exist_ok=True)
for name in ("alphas_2b.dat", "silicon.xml.sparseX.GAP_2020_10_15_180_13_48_39_9741", "alphas_3b.dat",
             "silicon.xml.sparseX.GAP_2020_10_15_180_13_48_39_9742", "corepot.dat", "compress.dat",
             "alphas_mb.dat"):
    with open(os.path.join(REF, "tests/stopping/1/gap_files", name)) as f:
        data = f.read()
    with open(os.path.join(HERE, "gap_files", name), "w") as f:
        f.write(data)

with open(os.path.join(REF, "tests/regression/cases/xrd_mad/expected/trajectory_out.xyz")) as f:
    lines = f.read().splitlines(keepends=True)
n = int(lines[0])
with open(os.path.join(HERE, "xrd_mad_frame0.xyz"), "w") as f:
    f.writelines(lines[: n + 2])
