"""Extract the count columns of the reference's saved outputs into small fixtures.

Reads (only in the build container, where the reference checkout exists):
  Tutorials/testdata/ps.dat  -- 32 P(k) bins, grid 64, kbin = kcenter = 1
                                (written by MEASUREMENTS.save_powerspectrum, pbi4/pbi4.py:555-563)
  Tutorials/testdata/bs.dat  -- 1156 triangles, grid 64, kbin = kcenter = 1, iopen = True
                                (written by MEASUREMENTS.save_bispectrum, pbi4/pbi4.py:574-594)

Writes (committed data, no reference source):
  ps_N64_kb1_kc1.txt : columns k, Nmodes, PSN exactly as saved (k and PSN as float32 text)
  bs_N64_kb1_kc1.txt : columns k1/kF k2/kF k3/kF N_tr (integers)

The P/B value columns are not extracted: the input catalogue they were measured on is not
shipped with the reference, so no implementation can reproduce them.
"""
import os
import sys

import numpy as np

REF = "/root/reference/Tutorials/testdata"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    if not os.path.isdir(REF):
        print("reference testdata not present; fixtures are already committed")
        return 0
    ps = np.loadtxt(os.path.join(REF, "ps.dat"))
    bs = np.loadtxt(os.path.join(REF, "bs.dat"))
    assert ps.shape == (32, 7) and bs.shape == (1156, 11)
    with open(os.path.join(HERE, "ps_N64_kb1_kc1.txt"), "w") as f:
        f.write("# from reference Tutorials/testdata/ps.dat: k  Nmodes  PSN (grid 64, kbin=kcenter=1)\n")
        for k, nm, psn in zip(ps[:, 0], ps[:, 5], ps[:, 6]):
            f.write("%.8e %d %.8e\n" % (k, int(nm), psn))
    with open(os.path.join(HERE, "bs_N64_kb1_kc1.txt"), "w") as f:
        f.write("# from reference Tutorials/testdata/bs.dat: k1/kF k2/kF k3/kF N_tr "
                "(grid 64, kbin=kcenter=1, iopen=True)\n")
        for row in bs:
            f.write("%d %d %d %d\n" % (int(row[0]), int(row[1]), int(row[2]), int(round(row[8]))))
    return 0


if __name__ == "__main__":
    sys.exit(main())
