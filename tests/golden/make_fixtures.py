"""Convert the reference's own fixture files into small .npz vectors.

Reads, as data only, from /root/reference/ArmandoList:
  out.txt        first neighbour-list dump of the 6462-particle dam break
                 (rows "ip j1..j21", MAXN = 21, self padded; sorted indices)
  old.txt        GPU output of the known-answer test check()
                 (armando2Dlist.cu:2451-2548): rows (drho/dt, dvx/dt, dvy/dt)
  reference.txt  analytic values of the same test, same row order
  new_pos.txt, new_vel.txt, new_info.txt
                 the 85,773-particle dam-break state written by outputData
                 (armando2Dcpu.cu:1542-1566): positions, velocities and
                 material (1 fluid, -1 wall) of every particle, float32 as
                 the C path holds them (its %+14.8e text round-trips float32)
Run once in a container where the reference is mounted; the .npz files are
committed and the tests only read them.
"""
import os
import sys

import numpy as np

SRC = "/root/reference/ArmandoList"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} not present")
    with open(os.path.join(SRC, "out.txt")) as f:
        lines = f.read().splitlines()
    assert lines[1].strip() == "Particles: 6462"
    rows = [list(map(int, ln.split())) for ln in lines[3:3 + 6462]]
    lst = np.array(rows, dtype=np.int32)
    assert lst.shape == (6462, 22) and np.array_equal(lst[:, 0], np.arange(6462))
    np.savez_compressed(os.path.join(HERE, "damlist6462.npz"), list=lst[:, 1:])
    old = np.loadtxt(os.path.join(SRC, "old.txt"), dtype=np.float64).astype(np.float32)
    ref = np.loadtxt(os.path.join(SRC, "reference.txt"), dtype=np.float64).astype(np.float32)
    assert old.shape == (10000, 3) and ref.shape == (10000, 3)
    np.savez_compressed(os.path.join(HERE, "kat_check.npz"), gpu=old, analytic=ref)
    pos = np.loadtxt(os.path.join(SRC, "new_pos.txt"), dtype=np.float64)
    vel = np.loadtxt(os.path.join(SRC, "new_vel.txt"), dtype=np.float64)
    info = np.loadtxt(os.path.join(SRC, "new_info.txt"), dtype=np.float64)
    assert pos.shape == (85773, 2) and vel.shape == (85773, 2) and info.shape == (85773, 3)
    p32, v32 = pos.astype(np.float32), vel.astype(np.float32)
    # lossless: the text is the float32 value printed with nine digits
    assert np.array_equal(np.char.mod("%+14.8e", p32.astype(np.float64)),
                          np.char.mod("%+14.8e", pos))
    assert np.all(np.isin(info[:, 0], (-1, 1)))
    assert np.all(info[:, 1] == info[0, 1]) and np.all(info[:, 2] == info[0, 2])
    np.savez_compressed(os.path.join(HERE, "damstate85773.npz"), pos=p32, vel=v32,
                        mat=info[:, 0].astype(np.int8),
                        mass_h=info[0, 1:].astype(np.float32))
    print("wrote damlist6462.npz, kat_check.npz, damstate85773.npz")


if __name__ == "__main__":
    main()
