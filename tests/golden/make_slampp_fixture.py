#!/usr/bin/env python3
"""Writes the SLAM++ conversion fixture pair from the reference's own data:
the first K odometry steps (and their landmark records) of
data/cityTrees10k.txt, and the matching delta / u / x / z lines of the
reference's shipped conversion data/cityTrees10k.log.txt.  Run in the build
container (it reads /root/reference); the outputs are committed."""
import os
import sys

REF = "/root/reference/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
K = int(sys.argv[1]) if len(sys.argv) > 1 else 300


def main():
    src, keep_ids, odo = [], {0}, 0
    for line in open(os.path.join(REF, "cityTrees10k.txt")):
        p = line.split()
        if not p:
            continue
        if p[0] == "ODOMETRY":
            if odo == K:
                continue
            odo += 1
            keep_ids.add(int(p[2]))
            src.append(line)
        elif p[0] == "LANDMARK" and int(p[1]) in keep_ids and odo <= K:
            src.append(line)
    log = []
    for line in open(os.path.join(REF, "cityTrees10k.log.txt")):
        p = line.split()
        if p[0] == "delta" or (p[0] in ("u", "x", "z") and int(p[1]) <= K):
            log.append(line)
    with open(os.path.join(OUT, "cityTrees10k_head.txt"), "w") as f:
        f.writelines(src)
    with open(os.path.join(OUT, "cityTrees10k_head.log.txt"), "w") as f:
        f.writelines(log)
    print(f"{len(src)} SLAM++ lines, {len(log)} log lines")


if __name__ == "__main__":
    main()
