"""Writes tests/golden/brick_trace.jsonl.gz: the reference brick topology and
region tracker (FluxFX_Source/native/bricks.hpp, regions.hpp, built in place
by oracle/build_ref.sh into oracle/_ref/) driven by oracle/ref/brick_trace.cpp's
seeded command sequence.  Needs /root/reference and g++."""
import gzip
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]


def main():
    subprocess.run(["bash", str(ROOT / "oracle" / "build_ref.sh")], check=True)
    out = subprocess.run([str(ROOT / "oracle" / "_ref" / "brick_trace")], check=True, capture_output=True).stdout
    with open(HERE / "brick_trace.jsonl.gz", "wb") as raw, \
            gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as f:
        f.write(out)
    print(len(out.splitlines()), "trace lines")


if __name__ == "__main__":
    main()
