"""Writes tests/golden/mesh_sdf_*.sdf: baked mesh-SDF data of the reference.

Copies two of the reference's baked mesh-SDF assets (Assets/sdf/*.sdf, the output of
Engine/DDGI/SDFBaker.cpp:158-200: a cereal binary archive of the size, the mip count and the R16F
texels of each level) as data files for the asset parser, the baker's mip rule and the
rasterization tests. Run in a container where /root/reference exists; the copies are committed.
"""
import os
import shutil

SRC = "/root/reference/Assets/sdf"
HERE = os.path.dirname(os.path.abspath(__file__))
COPIES = {"BOX_Black_0.001_1673383391963.sdf": "mesh_sdf_box_32x32x35.sdf",
          "BOX_Black_0.002_1673383391976.sdf": "mesh_sdf_box_51x32x32.sdf"}


def main():
    for src, dst in COPIES.items():
        shutil.copyfile(os.path.join(SRC, src), os.path.join(HERE, dst))
    print("copied", len(COPIES), "assets")


if __name__ == "__main__":
    main()
