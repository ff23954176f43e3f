"""Golden restart files written by the reference's own write_restart_files.

Runs in the build container only (it needs /root/reference and ROCm flang):

    make -C oracle ref
    python tests/golden/make_restart_fixtures.py

oracle/_ref/ref_restart_driver fills the reference's module variables with
seeded values and calls write_restart_files (src/write_restart_files.f90:22-47);
the inputs and the bytes of the file it wrote are stored in restart_ref.npz.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_restart_driver")
FIELDS = ("flmass", "u", "v", "w", "p", "te", "ed", "t", "vis", "uu", "vv", "ww", "uv", "uw", "vw",
          "uo", "vo", "wo", "teo", "edo")
CELL_ONLY = ("uu", "vv", "ww", "uv", "uw", "vw")


def case(tag, nc, nif, ntot, const_mflux, itime, seed):
    rng = np.random.default_rng(seed)
    time = float(rng.random() * 10)
    gradpcmf = float(rng.standard_normal())
    size = {k: (nif if k == "flmass" else nc if k in CELL_ONLY else ntot) for k in FIELDS}
    fields = {k: rng.standard_normal(size[k]) for k in FIELDS}
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "restart_in.txt"), "w") as f:
            f.write(f"{nc} {nif} {ntot} {int(const_mflux)} {itime}\n")
        with open(os.path.join(d, "restart_in.bin"), "wb") as f:
            f.write(np.array([time, gradpcmf]).tobytes())
            for k in FIELDS:
                f.write(fields[k].tobytes())
        subprocess.run([DRIVER], cwd=d, check=True, capture_output=True)
        data = open(os.path.join(d, "restart_ref"), "rb").read()
    out = {f"{tag}_{k}": v for k, v in fields.items()}
    out.update({f"{tag}_file": np.frombuffer(data, np.uint8).copy(),
                f"{tag}_meta": np.array([nc, nif, ntot, int(const_mflux), itime], np.int64),
                f"{tag}_scal": np.array([time, gradpcmf])})
    return out


def main():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    data = {}
    data.update(case("plain", nc=7, nif=12, ntot=10, const_mflux=False, itime=42, seed=1))
    data.update(case("mflux", nc=5, nif=8, ntot=9, const_mflux=True, itime=7, seed=2))
    path = os.path.join(HERE, "restart_ref.npz")
    np.savez_compressed(path, **data)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
