"""Write tests/golden/toy_0.{col,dst,vlabel,g,pangolin.txt}: the reference's
converters (converter/toGAMMA.cpp, toGSI.cpp, toPangolin.cpp, compiled
unmodified from their sources by `make -C oracle ref`) run on the preprocessed
toy graph tests/golden/toy_0.pre.bin.  Needs the reference checkout, so it runs
in the build container only."""
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as td:
        g = os.path.join(td, "toy_0.bin")
        shutil.copy(os.path.join(HERE, "toy_0.pre.bin"), g)
        for tool in ("toGAMMA", "toGSI", "toPangolin"):
            subprocess.run([os.path.join(ROOT, "oracle", "_ref", tool), g], check=True, capture_output=True)
        for ext in ("col", "dst", "vlabel", "g", "pangolin.txt"):
            shutil.copy(os.path.join(td, "toy_0." + ext), os.path.join(HERE, "toy_0." + ext))
    print("wrote the converter fixtures of toy_0")


if __name__ == "__main__":
    main()
