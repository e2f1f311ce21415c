"""Writes tests/golden/data/MANIFEST.json: name, shape, dtype, sha256 and the
reference test line that pins each fixture."""
import hashlib
import json
import os
import sys

import numpy as np

# fixture name -> (source file in the reference, test that asserts on it)
SOURCES = {
    "randconf": ("test/data/O3.jld", "test/tests_O3.jl:3,51"),
    "afterlocal_hsfield": ("test/data/O3.jld", "test/tests_O3.jl:192"),
    "greens": ("test/data/O3.jld", "test/tests_O3.jl:223,248"),
    "Ul": ("test/data/O3.jld", "test/tests_O3.jl:220"),
    "Dl": ("test/data/O3.jld", "test/tests_O3.jl:221"),
    "Tl": ("test/data/O3.jld", "test/tests_O3.jl:222"),
    "Ur": ("test/data/O3.jld", "test/tests_O3.jl:217"),
    "Dr": ("test/data/O3.jld", "test/tests_O3.jl:218"),
    "Tr": ("test/data/O3.jld", "test/tests_O3.jl:219"),
    "delta_i": ("test/data/O3.jld", "test/tests_O3.jl:182"),
    "M": ("test/data/O3.jld", "test/tests_O3.jl:183"),
    "afterupdate_greens": ("test/data/O3.jld", "test/tests_O3.jl:186"),
    "afterlocal_greens": ("test/data/O3.jld", "test/tests_O3.jl:191"),
    "B_QR_U": ("test/data/linalg.jld", "test/tests_linalg.jl:107-112"),
    "B_QR_D": ("test/data/linalg.jld", "test/tests_linalg.jl:107-112"),
    "B_QR_T": ("test/data/linalg.jld", "test/tests_linalg.jl:107-112"),
}


def main(out):
    entries = []
    for fn in sorted(os.listdir(out)):
        if not fn.endswith(".npy"):
            continue
        name = fn[:-4]
        a = np.load(os.path.join(out, fn), allow_pickle=False)
        with open(os.path.join(out, fn), "rb") as f:
            digest = hashlib.sha256(f.read()).hexdigest()
        src, test = SOURCES.get(name, ("test/data/O3.jld", "test/tests_O3.jl:62-175,298-304"))
        entries.append({"name": name, "shape": list(a.shape), "dtype": str(a.dtype),
                        "sha256": digest, "source": src, "pinned_by": test})
    with open(os.path.join(out, "MANIFEST.json"), "w") as f:
        json.dump(entries, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1])
