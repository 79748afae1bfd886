"""Write tests/golden/toy_*: the reference's offline tools on small graphs.

Runs oracle/_ref/ref_preprocess and oracle/_ref/ref_kcore (the reference's
kmeans/preprocess.cpp and kmeans/kcore.cpp, compiled unmodified from their
sources by `make -C oracle ref`) on seeded SNAP-style edge lists with
duplicate edges, self loops, an id offset and isolated ids, and keeps the
inputs and outputs:
  toy_<k>.txt              input edge list
  toy_<k>.pre.bin          preprocess output
  toy_<k>.kcore.bin        kcore -nb 4 output (recoded ids)
  toy_<k>.kcore.vbmap      its view-bin map (<bin>.2.hop.vbmap)
Needs the reference checkout, so it runs in the build container only.
"""
import os
import random
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def write_txt(path, seed, n, m, offset):
    rng = random.Random(seed)
    lines = ["# toy graph seed %d" % seed]
    # power-law-ish endpoints, plus duplicates and self loops
    w = [1.0 / (i + 1) ** 0.7 for i in range(n)]
    for _ in range(m):
        a = rng.choices(range(n), w)[0]
        b = rng.choices(range(n), w)[0]
        lines.append("%d\t%d" % (a + offset, b + offset))
    for _ in range(m // 20):
        a = rng.randrange(n)
        lines.append("%d %d" % (a + offset, a + offset))
    lines.append("%d %d" % (offset, offset + n + 5))  # ids with no other edges
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    pre = os.path.join(ROOT, "oracle", "_ref", "ref_preprocess")
    kc = os.path.join(ROOT, "oracle", "_ref", "ref_kcore")
    specs = [(11, 60, 180, 3), (12, 300, 2400, 1), (13, 1000, 9000, 0)]
    for k, (seed, n, m, off) in enumerate(specs):
        txt = os.path.join(HERE, "toy_%d.txt" % k)
        write_txt(txt, seed, n, m, off)
        with tempfile.TemporaryDirectory() as td:
            t = os.path.join(td, "g.txt")
            shutil.copy(txt, t)
            subprocess.run([pre, "-f", t], check=True, capture_output=True)
            shutil.copy(os.path.join(td, "g.bin"), os.path.join(HERE, "toy_%d.pre.bin" % k))
            subprocess.run([kc, "-f", os.path.join(td, "g.bin"), "-nb", "4"], check=True,
                           capture_output=True)
            shutil.copy(os.path.join(td, "g.bin"), os.path.join(HERE, "toy_%d.kcore.bin" % k))
            shutil.copy(os.path.join(td, "g.bin.2.hop.vbmap"),
                        os.path.join(HERE, "toy_%d.kcore.vbmap" % k))
    print("wrote", len(specs), "toy graph fixtures")


if __name__ == "__main__":
    main()
