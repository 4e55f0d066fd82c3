"""Copies the reference's own multi-patch topology test inputs (data files of
TestASMs3D.Connect and TestASMs3D.ConnectUneven, src/ASM/Test/ASM3DTests.h:
96-125) into tests/golden/refdata, unchanged, so that the tests can read
them where the reference tree is absent.  Run from the repository root."""
import glob
import os
import shutil

SRC = "/root/reference/src/ASM/Test/refdata"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "refdata")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    names = [os.path.basename(p) for p in glob.glob(os.path.join(SRC, "3d_uneven*"))]
    for k in range(8):
        names += [f"DomainDecomposition_MPI_3D_4_orient{k}.xinp", f"cube-8-orient{k}.g2"]
    for n in sorted(names):
        shutil.copyfile(os.path.join(SRC, n), os.path.join(DST, n))
    print(len(names), "files")
