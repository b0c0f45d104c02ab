"""Writes tests/golden/sphere_pairs_2_1024.json from the reference's CPU
broadphase dump physicboxSap2.16_Fermi/cudaPhysics/2_1024.txt.

That file was produced by DemoParticles::cpuBroadphase
(cudaPhysics/demoParticles.h:1496-1568, dump format :1539-1557): per frame, every
directed overlapping pair "(i j)" followed by the (x y z r) of sphere i, of
sphere j and of a debug sphere, then "overlap pairs N", a rule, the pairEntries
row starts and "pairsEntries N".  Positions are printed with 6 significant
digits.  The fixture keeps, per frame: the directed pairs in file order, the
positions of every sphere that appears, and the entries list -- data only.

Run in the build container (reads /root/reference); the JSON is committed.
"""
import json
import os
import re
import sys

SRC = "/root/reference/physicboxSap2.16_Fermi/cudaPhysics/2_1024.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere_pairs_2_1024.json")


def parse(text):
    frames = []
    cur = {"pairs": [], "pos": {}, "entries": []}
    lines = text.splitlines()
    i = 0
    pair_re = re.compile(r"\((\d+)\s+(\d+)\)")
    while i < len(lines):
        line = lines[i].strip()
        m = re.match(r"pairsEntries\s+(\d+)(.*)$", line)
        if m:
            cur["n_entries_reported"] = int(m.group(1))
            frames.append(cur)
            cur = {"pairs": [], "pos": {}, "entries": []}
            rest = m.group(2).strip()
            if rest:
                lines[i] = rest
                continue
            i += 1
            continue
        m = pair_re.match(line)
        if m:
            a, b = int(m.group(1)), int(m.group(2))
            pa = [float(v) for v in lines[i + 1].split()]
            pb = [float(v) for v in lines[i + 2].split()]
            cur["pairs"].append([a, b])
            cur["pos"][str(a)] = pa
            cur["pos"][str(b)] = pb
            i += 4
            continue
        m = re.match(r"overlap pairs\s+(\d+)", line)
        if m:
            cur["n_pairs_reported"] = int(m.group(1))
            i += 1
            continue
        if line.startswith("---") or not line:
            i += 1
            continue
        if re.fullmatch(r"\d+", line):
            cur["entries"].append(int(line))
            i += 1
            continue
        raise ValueError(f"unparsed line {i}: {line!r}")
    if cur["pairs"]:
        frames.append(cur)
    return frames


def main():
    with open(SRC) as f:
        frames = parse(f.read())
    with open(OUT, "w") as f:
        json.dump({"source": "physicboxSap2.16_Fermi/cudaPhysics/2_1024.txt", "frames": frames}, f,
                  separators=(",", ":"))
    print(f"{len(frames)} frames, {sum(len(fr['pairs']) for fr in frames)} directed pairs -> {OUT}")


if __name__ == "__main__":
    sys.exit(main())
