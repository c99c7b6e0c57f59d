"""Extract compact golden fixtures from the reference's shipped run outputs.

Run in a container where the reference tree is mounted (it is read here and
nowhere else); the outputs are data files committed next to this script:

  v251_run1_input.txt     input.txt of v25/v251/run1 (parameters of the run)
  v251_run1_output.txt    the printed observable lines of output.dat
  v251_run1_conf.txt      conf.dat after step 100 (state + cluster labels)
  v251_run1_compConf.txt  neighbour-list reference positions (12 decimals)
  v251_run1_iniPos.txt    ini_pos lattice (default stream precision)
  v251_run1_posvel.txt    pos_vel_<n>.dat, n = 0..50, concatenated with
                          "# file n" markers (25 particles every 2 steps)
  v251_run1_correlation.txt.gz
                          correlation.dat (pair histograms of steps 0, 10,
                          ..., 100, normalised), gzip-compressed
  v251_run1_magmom.txt    magmom.dat (total dipole moment after every step)
  v251_run1_counters.txt  counters.dat after step 100 (sums over the steps)
  v25_run1_nve.txt        time, K and U columns of v25/run1/output.dat for the
                          two NVE segments used as drift statistics
  v251_run1_ptens.txt     ptens.dat of v25/v251/run1 (time, Helfand moment
                          sum r_a p_b, total pressure tensor; every step)
  v25_input.txt           v25/input.txt (two species of 50 short rods, two
                          off-centre dipoles each, K=10, Signal=0)
  v25_output_seg.txt      v25/output.dat lines 1220-1880: the 661 printed
                          steps of one uninterrupted run of v25/input.txt
                          (time and the 11 observables; the two timing
                          columns are dropped)

All of these are outputs the reference itself wrote; nothing here is source.
"""
import gzip
import os
import sys

REF = "/root/reference/v25"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not mounted; fixtures are already committed")
    run = os.path.join(REF, "v251", "run1")
    with open(os.path.join(run, "input.txt")) as f:
        open(os.path.join(HERE, "v251_run1_input.txt"), "w").write(f.read())
    # output.dat holds two appended copies of the same 101-step run; keep the
    # first block of per-print lines (steps 0, 10, ..., 100)
    out = [l for l in open(os.path.join(run, "output.dat"))
           if l.count("|") == 1 and l[:1].isdigit() and ". " in l][:11]
    open(os.path.join(HERE, "v251_run1_output.txt"), "w").writelines(out)
    for src, dst in [("conf.dat", "v251_run1_conf.txt"), ("compConf.dat", "v251_run1_compConf.txt"),
                     ("iniPos.dat", "v251_run1_iniPos.txt"), ("magmom.dat", "v251_run1_magmom.txt"),
                     ("counters.dat", "v251_run1_counters.txt"), ("ptens.dat", "v251_run1_ptens.txt")]:
        open(os.path.join(HERE, dst), "w").write(open(os.path.join(run, src)).read())
    with open(os.path.join(run, "correlation.dat"), "rb") as f:
        data = f.read()
    with gzip.GzipFile(os.path.join(HERE, "v251_run1_correlation.txt.gz"), "wb", mtime=0) as f:
        f.write(data)
    with open(os.path.join(HERE, "v251_run1_posvel.txt"), "w") as f:
        for n in range(51):
            f.write("# file %d\n" % n)
            f.write(open(os.path.join(run, "pos_vel", "pos_vel_%d.dat" % n)).read())
    open(os.path.join(HERE, "v25_input.txt"), "w").write(open(os.path.join(REF, "input.txt")).read())
    lines = open(os.path.join(REF, "output.dat")).read().splitlines()[1219:1880]
    with open(os.path.join(HERE, "v25_output_seg.txt"), "w") as f:
        for line in lines:
            head, tail = line.split("|")
            f.write("%s | %s\n" % (head.split()[0], " ".join(tail.split())))
    # v25/run1: N=100 dipolar spheres, Q=1e12, gamma=0 (NVE).  The segments
    # are the runs bounded by restarts of the time column.
    rows = []
    for line in open(os.path.join(REF, "run1", "output.dat")):
        if line.count("|") != 1 or not line[:1].isdigit():
            rows.append(None)
            continue
        head, tail = line.split("|")
        time = float(head.split()[0].rstrip("."))
        vals = tail.split()
        rows.append((time, float(vals[0]), float(vals[1])))
    with open(os.path.join(HERE, "v25_run1_nve.txt"), "w") as f:
        f.write("# line time K U  (v25/run1/output.dat, per-particle values)\n")
        for i, r in enumerate(rows):
            if r is not None and (4089 <= i + 1 <= 6089 or 6133 <= i + 1 <= 7394):
                f.write("%d %.12f %.12f %.12f\n" % (i + 1, r[0], r[1], r[2]))


if __name__ == "__main__":
    main()
