"""Copies two data files of the reference's own tests into tests/golden/ (run
once, in a tree that has the reference; the copies are what the tests read):

  * silicon.gap: tests/stopping/1/gap_files/silicon.gap, a potential file with
    distance_2b, angle_3b, soap_turbo (compressed) and core_pot blocks;
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

with open(os.path.join(REF, "tests/regression/cases/xrd_mad/expected/trajectory_out.xyz")) as f:
    lines = f.read().splitlines(keepends=True)
n = int(lines[0])
with open(os.path.join(HERE, "xrd_mad_frame0.xyz"), "w") as f:
    f.writelines(lines[: n + 2])
