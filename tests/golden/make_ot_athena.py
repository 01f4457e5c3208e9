"""Extract the Athena++ Orszag-Tang reference (500x500, t = 0.5) held by the
reference's validation suite (tests/validation/ot_reference.h5, checksum guarded by
tests/validation/test_validation.py:45-76) into tests/golden/ot_athena500_t05.npz:
rho and thermal P of its single snapshot plus its time.  The file is read with this
package's HDF5 reader, so the extraction also exercises that reader on a file another
writer produced.

The file is in the snapshot layout the reference's tests read (one group per field
with SNAPSHOT<k> datasets, /parameters/{M_s,beta,gamma}, /time/sim_time), so its
structure -- `h5dump -H -p` without the file name and the per-file object offsets --
is also written, to tests/golden/ot_athena500_layout.txt: the layout this package's
writer must reproduce for a one-snapshot 500 x 500 file.

    python tests/golden/make_ot_athena.py            # both files
    python tests/golden/make_ot_athena.py --layout   # the layout file only
"""
from __future__ import annotations

import hashlib
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from chimera_amd import h5io  # noqa: E402

SRC = Path("/root/reference/tests/validation/ot_reference.h5")
SHA256 = "45e66217f6a81bae0831f8201b0cf6b10cdd75df3b5ecfd24c1e7899b5e92b5d"
OUT = Path(__file__).resolve().parent / "ot_athena500_t05.npz"
LAYOUT = Path(__file__).resolve().parent / "ot_athena500_layout.txt"
H5DUMP = Path("/opt/conda/bin/h5dump")


def h5_layout(path) -> str:
    """`h5dump -H -p` of an HDF5 file with its name replaced by FILE and the OFFSET
    lines (where each writer placed the raw data) dropped."""
    out = subprocess.run([str(H5DUMP), "-H", "-p", str(path)], capture_output=True, text=True,
                         check=True).stdout
    out = out.replace(str(path), "FILE")
    return "".join(line + "\n" for line in out.splitlines() if "OFFSET" not in line)


def main() -> None:
    digest = hashlib.sha256(SRC.read_bytes()).hexdigest()
    if digest != SHA256:
        raise SystemExit(f"{SRC} checksum {digest} != {SHA256}")
    LAYOUT.write_text(h5_layout(SRC))
    print(f"wrote {LAYOUT}")
    if "--layout" in sys.argv[1:]:
        return
    d = h5io.read(str(SRC))
    assert len(d["rho"]) == 1 and d["rho"][0].shape == (500, 500)
    np.savez_compressed(OUT, rho=d["rho"][0], P=d["P"][0], time=d["time"],
                        source_sha256=np.array(SHA256))
    print(f"wrote {OUT} (t = {d['time'][0]})")


if __name__ == "__main__":
    main()
