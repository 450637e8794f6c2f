"""Copies the reference's own f64 test fixtures for the reconstruction step into this folder,
gzip-compressed. The files are data (inputs and expected outputs written by the reference's
`refs/f64/*.py` scripts and read by its `ptyrs/tests/engine_*.rs` tests); nothing of the
reference's code is copied. Run once in a container that has the reference mounted:

    python tests/golden/collect_fixtures.py /root/reference
"""

import gzip
import pathlib
import sys

FAMILIES = ("rpie_near_", "propagator_", "pynx_near_", "lsqml_near_", "lsqml_multislice_",
            "ml_grad")

def main(root):
    src = pathlib.Path(root) / "ptyrs" / "tests" / "engine" / "data"
    dst = pathlib.Path(__file__).resolve().parent
    for f in sorted(src.iterdir()):
        if f.name.startswith(FAMILIES) and f.suffix == ".txt":
            with gzip.open(dst / (f.name + ".gz"), "wb", compresslevel=9) as out:
                out.write(f.read_bytes())
            print(f.name)

if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
