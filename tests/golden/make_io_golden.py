"""Writes the restart-primitive and raw-export fixtures under tests/golden/io
with the reference's own routines (oracle/_ref, flang-built from the
reference's sources; this container only).

  gf_input.npy          a 5x4x3 array of mixed magnitudes, signs and special
                        values (zeros, subnormal, extremes)
  gf_reference.txt      export(grid_field) of it
                        (code/generated_code/grid_field/grid_field.f90:108-139)
  raw_nodes.npz         node arrays (with ghosts) of a 4 x 5 (Roberts) x 3 grid
  raw_input_<DL>.npy    a field of each data location on that grid
  raw_<DL>_pad<p>/      export_raw(m, x, dir, 'U', pad) of it, written by
                        export_3D_1C_steady (code/handwritten/IO/IO_export.f90:265-277)

Run from the repository root:  python tests/golden/make_io_golden.py
"""
import os
import shutil
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "io")
DL = {"CC": 0, "Fx": 1, "Fy": 2, "Fz": 4, "Ex": 6, "Ey": 5, "Ez": 3, "N": 7}
RAW_CASES = [("CC", 1), ("CC", 0), ("Fx", 0), ("Ey", 1), ("N", 0)]


def gf_input():
    rng = np.random.default_rng(2024)
    f = rng.standard_normal((5, 4, 3)) * 10.0 ** rng.integers(-25, 25, (5, 4, 3))
    flat = f.reshape(-1, order="F")
    flat[:8] = [0.0, -0.0, 5e-324, 1.7976931348623157e308, 0.1, 0.095, 9.5e14, -949999999999999.9]
    return flat.reshape(f.shape, order="F")


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    os.makedirs(OUT, exist_ok=True)
    f = gf_input()
    np.save(os.path.join(OUT, "gf_input.npy"), f)
    O.ref_gf_export(os.path.join(OUT, "gf_reference.txt"), f)
    hn = [O.uniform_nodes(0.0, 1.0, 4), O.roberts_both_nodes(-1.0, 1.0, 5, 1.05), O.uniform_nodes(0.0, 2.0, 3)]
    np.savez(os.path.join(OUT, "raw_nodes.npz"), hx=hn[0], hy=hn[1], hz=hn[2])
    rng = np.random.default_rng(7)
    for name, code in DL.items():
        shape = tuple(len(h) if (code >> i) & 1 else len(h) - 1 for i, h in enumerate(hn))
        np.save(os.path.join(OUT, f"raw_input_{name}.npy"), rng.standard_normal(shape) * 10.0 ** rng.integers(-3, 3, shape))
    for name, pad in RAW_CASES:
        d = os.path.join(OUT, f"raw_{name}_pad{pad}")
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        f = np.load(os.path.join(OUT, f"raw_input_{name}.npy"))
        O.ref_export_raw(hn, DL[name], f, d + "/", "U", pad)
    print("fixtures written under", OUT)


if __name__ == "__main__":
    main()
