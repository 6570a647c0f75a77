"""Copies the input files of the reference's 2D benchmarks (TOML, entity
sets, particle positions; data only) from a reference checkout into this
folder, one subfolder per benchmark.  Usage:
python tests/golden/benchmarks/copy_fixtures.py /root/reference"""
import os
import shutil
import sys

NAMES = ["uniaxial_stress", "uniaxial_nodal_forces", "uniaxial_particle_traction"]


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for name in NAMES:
        src = os.path.join(ref, "benchmarks", "2d", name)
        dst = os.path.join(here, name)
        os.makedirs(dst, exist_ok=True)
        for f in sorted(os.listdir(src)):
            if f.endswith((".toml", ".json")):
                shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
                print(os.path.join(name, f))


if __name__ == "__main__":
    main(sys.argv[1])
