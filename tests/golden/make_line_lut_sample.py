"""Write tests/golden/line_lut_sample.npz: a sampled copy of the reference's tabulated
line-integrated kernels (Partia.jl/src/julia/KernelInterpolation/table/
line_integrated_kernel_tables.jl), read as text.  For each kernel it keeps every 32nd
entry plus the last 4 of Q, I and Inorm.  The fixture pins the regenerated tables of the
oracle and of the HIP library; the full table is not copied.

Run in a container that has the reference mounted:
    python tests/golden/make_line_lut_sample.py [/root/reference]
"""
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
PATH = os.path.join(REF, "Partia.jl/src/julia/KernelInterpolation/table/line_integrated_kernel_tables.jl")
KERNELS = ["M4_spline", "M5_spline", "M6_spline", "C2_Wendland", "C4_Wendland", "C6_Wendland"]


def main():
    txt = open(PATH).read()
    tabs = {}
    for m in re.finditer(r"const (_\w+) = SVector\{4096,Float64\}\((.*?)\)", txt, re.S):
        tabs[m.group(1)] = np.array([float(v) for v in m.group(2).replace("\n", "").split(",") if v.strip()])
    idx = np.unique(np.concatenate([np.arange(0, 4096, 32), np.arange(4092, 4096)]))
    out = {"index": idx}
    for k in KERNELS:
        for part in ("Q", "I", "Inorm"):
            out[f"{k}_{part}"] = tabs[f"_{k}_{part}"][idx]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "line_lut_sample.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, len(idx), "entries per table")


if __name__ == "__main__":
    main()
